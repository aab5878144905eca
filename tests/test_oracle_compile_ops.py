"""Pins of the oracle's compile step, gather apply, dot and CG, and the product compile step's
refusals (CPU only).

- DOF numbering: the reference's constraint goldens (tests/reference/data/constraints, OpenSees
  955ceb0 equation numbers per node): the equation classes, the fixed set and the equation
  count must equal OpenSees' on every valid case, also after a node renumbering; every invalid
  case must be refused with the reference's pattern (tests/unit/test_dof_numbering.cpp:129-175).
- RCM / element order: the reference's own properties (tests/unit/test_renumber.cpp:57-129) on
  the same scrambled grids.
- Gather apply: against K assembled densely from element matrices that the Brick goldens pin.
- Dot: exact sums, and sums whose rounding depends on the chunk and lane order, computed by
  hand from the order's definition (vector_ops.cpp:19-124).
- CG + Jacobi: against a dense direct solve.
"""
from fractions import Fraction

import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from oracle import fxo
from tests import meshes

GOLD = np.load(__file__.replace("test_oracle_compile_ops.py", "golden/constraints.npz"))
CASES = sorted({k.split("/")[0] for k in GOLD.keys()})


def case_model(name):
    """A constraint case as a model dict: nodes sorted by tag (node i has tag i + 1 in the
    compile step's rules), no elements."""
    nodes = GOLD[name + "/nodes"]
    order = np.argsort(nodes[:, 0], kind="stable")
    tags = nodes[order, 0].astype(int)
    index = {t: i for i, t in enumerate(tags)}
    assert list(tags) == list(range(1, len(tags) + 1))
    fix = np.array([[index[t], d - 1] for t, d in GOLD[name + "/fix"]], np.int32).reshape(-1, 2)
    eq = GOLD[name + "/equal"]
    tie = np.array([[index[r], index[c], d - 1] for r, c, d, _ in eq], np.int32).reshape(-1, 3)
    return {"nnode": len(tags), "xyz": nodes[order, 2:5].copy(), "nelem": 0, "conn": np.zeros((0, 8), np.int32),
            "group": np.zeros(0, np.int32), "fix": fix, "tie": tie,
            "tie_command": eq[:, 3].astype(np.int32).copy() if len(eq) else None}, tags


def opensees_classes(name):
    os_eq = {}
    for row in GOLD[name + "/opensees"]:
        for d in range(int(row[1])):
            os_eq[(int(row[0]), d)] = int(row[2 + d])
    return os_eq


def disagreements(node_eq, node_rank, tags, os_eq):
    a, b, bad = {}, {}, 0
    for i, t in enumerate(tags):
        for d in range(3):
            e = int(node_eq[3 * node_rank[i] + d])
            o = os_eq[(int(t), d)]
            if (e < 0) != (o < 0):
                bad += 1
                continue
            if e < 0:
                continue
            if a.setdefault(e, o) != o or b.setdefault(o, e) != e:
                bad += 1
    return bad


def test_constraint_goldens_present():
    assert len(CASES) == 11
    assert sum(str(GOLD[c + "/expect"]) == "valid" for c in CASES) == 6


@pytest.mark.parametrize("name", CASES)
def test_dof_numbering_matches_opensees(name):
    expect = str(GOLD[name + "/expect"])
    model, tags = case_model(name)
    if expect == "valid":
        os_eq = opensees_classes(name)
        os_neq = max(os_eq.values()) + 1
        om = fxo.OracleMesh(model, renumber=False, coloring=False)
        assert disagreements(om.node_eq, om.node_rank, tags, os_eq) == 0
        assert om.neq == os_neq
        pm = core.Mesh(model, renumber=False, coloring=False)
        np.testing.assert_array_equal(pm.node_eq, om.node_eq)
        # a renumbering changes equation numbers, never the classes
        rng = S.Lcg(3)
        perm = list(range(len(tags)))
        for i in range(len(perm), 1, -1):
            j = int((rng.symmetric() + 1) * 0.5 * i) % i
            perm[i - 1], perm[j] = perm[j], perm[i - 1]
        rank = np.empty(len(tags), np.int64)
        rank[np.array(perm)] = np.arange(len(tags))
        # renumbered equations: number the union-find classes in the permuted order
        eq_perm = np.full(3 * len(tags), -1)
        seen, nxt = {}, 0
        for internal in range(len(tags)):
            m = perm[internal]
            for d in range(3):
                e = om.node_eq[3 * om.node_rank[m] + d]
                if e < 0:
                    continue
                if e not in seen:
                    seen[e] = nxt
                    nxt += 1
                eq_perm[3 * internal + d] = seen[e]
        assert disagreements(eq_perm, rank, tags, os_eq) == 0
    else:
        words = expect.split(" ", 1)[1]
        with pytest.raises(fxo.RefusedModel) as info:
            fxo.OracleMesh(model, renumber=False, coloring=False)
        assert fxo.REFUSALS[info.value.code] in words or words in fxo.REFUSALS[info.value.code] or (
            info.value.code == 2 and "equalDOF commands" in words), (words, info.value.code)
        with pytest.raises(core.FxError) as perr:
            core.Mesh(model, renumber=False, coloring=False)
        assert words in str(perr.value), (words, str(perr.value))


def scrambled_grid(nx, ny, nz, scramble=True):
    """tests/unit/test_renumber.cpp:19-57: node tags permuted by the LCG (seed 9)."""
    n = (nx + 1) * (ny + 1) * (nz + 1)
    tag_of = list(range(1, n + 1))
    if scramble:
        rng = S.Lcg(9)
        for i in range(n - 1, 0, -1):
            j = int((rng.symmetric() + 1) * 0.5 * (i + 1)) % (i + 1)
            tag_of[i], tag_of[j] = tag_of[j], tag_of[i]
    by_tag = [0] * n
    for v in range(n):
        by_tag[tag_of[v] - 1] = v
    xyz = np.zeros((n, 3))
    for t in range(n):
        v = by_tag[t]
        xyz[t] = (v % (nx + 1), (v // (nx + 1)) % (ny + 1), v // ((nx + 1) * (ny + 1)))
    vid = lambda i, j, k: (k * (ny + 1) + j) * (nx + 1) + i  # noqa: E731
    conn = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = [vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k), vid(i, j, k + 1),
                     vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1), vid(i, j + 1, k + 1)]
                conn.append([tag_of[a] - 1 for a in c])
    conn = np.array(conn, np.int32)
    return {"nnode": n, "xyz": xyz, "nelem": len(conn), "conn": conn, "group": np.zeros(len(conn), np.int32)}


def test_rcm_off_keeps_file_order():
    om = fxo.OracleMesh(scrambled_grid(3, 3, 4), renumber=False, coloring=False)
    np.testing.assert_array_equal(om.node_order, np.arange(om.nnode))
    np.testing.assert_array_equal(om.elem_order, np.arange(om.nelem))


def test_rcm_is_a_valid_deterministic_permutation():
    m = scrambled_grid(4, 3, 6)
    a = fxo.OracleMesh(m, coloring=False)
    b = fxo.OracleMesh(m, coloring=False)
    np.testing.assert_array_equal(np.sort(a.node_order), np.arange(a.nnode))
    np.testing.assert_array_equal(np.sort(a.elem_order), np.arange(a.nelem))
    np.testing.assert_array_equal(a.node_order, b.node_order)
    np.testing.assert_array_equal(a.node_rank[a.node_order], np.arange(a.nnode))


def test_rcm_recovers_a_small_bandwidth():
    m = scrambled_grid(6, 5, 20)
    off = fxo.OracleMesh(m, renumber=False, coloring=False)
    rcm = fxo.OracleMesh(m, coloring=False)
    bw_off, bw_rcm = off.bandwidth(), rcm.bandwidth()
    section = 7 * 6
    assert bw_off > 500
    assert bw_rcm <= 5 * section // 2
    assert bw_rcm * 5 < bw_off


def test_elements_follow_their_nodes():
    m = scrambled_grid(5, 4, 8)
    om = fxo.OracleMesh(m, coloring=False)
    prev = -1
    for e in om.elem_order:
        k = min(om.node_rank[m["conn"][e]])
        assert k >= prev
        prev = k


def test_rcm_on_a_chain_of_bricks_has_the_minimal_bandwidth():
    """A 1 x 1 x n column: BFS levels are whole 4-node cross-sections, so RCM reaches the
    smallest bandwidth of the chain, 7 (an element spans two consecutive levels)."""
    m = scrambled_grid(1, 1, 12)
    om = fxo.OracleMesh(m, coloring=False)
    assert om.bandwidth() == 7


def assembled(model, params, om):
    """Dense K from element matrices (Brick goldens pin fxo_brick_force_elastic's K)."""
    L = fxo.lib()
    K = np.zeros((om.neq, om.neq))
    xyz = model["xyz"]
    for e in range(model["nelem"]):
        conn = model["conn"][e]
        p = fxo.elastic_params([params[model["group"][e]]])[0]
        xe = np.ascontiguousarray(xyz[conn], np.float64)
        Ke = np.zeros(576)
        L.fxo_brick_force_elastic(p, xe, np.zeros(24), np.zeros(24), np.zeros(24), Ke.ctypes.data)
        Ke = Ke.reshape(24, 24)
        eqs = [om.node_eq[3 * om.node_rank[n] + d] for n in conn for d in range(3)]
        for a, ea in enumerate(eqs):
            if ea < 0:
                continue
            for b, eb in enumerate(eqs):
                if eb >= 0:
                    K[ea, eb] += Ke[a, b]
    return K


@pytest.mark.parametrize("name", ["distorted3", "column6_tied"])
def test_gather_apply_equals_assembled_matrix(name):
    model = meshes.distorted_box(3) if name == "distorted3" else meshes.column(6)
    for p in model["params"]:
        p["b"] = (0.0, 0.0, 0.0)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])  # parameter sets in compiled order
    K = assembled(model, model["params"], om)
    x = meshes.random_vector(om.neq, 4)
    y = om.apply(op, x)
    assert np.max(np.abs(y - K @ x)) <= 1e-12 * np.max(np.abs(K @ x))
    d = om.diagonal(op)
    assert np.max(np.abs(d - np.diag(K))) <= 1e-12 * np.max(np.abs(np.diag(K)))


def test_dot_exact_sums():
    rng = np.random.default_rng(5)
    for n in (1, 7, 2048, 2049, 10000):
        x = rng.integers(-2**20, 2**20, n).astype(np.float64) / 1024.0
        y = rng.integers(-2**10, 2**10, n).astype(np.float64)
        want = sum(Fraction(a) * Fraction(b) for a, b in zip(x, y))
        for lanes in (1, 64):
            assert Fraction(fxo.dot(x, y, lanes)) == want


def test_dot_order_within_and_across_chunks():
    """With lanes = 1 a chunk is summed left to right; with 64 lanes, lane l sums entries
    l, l + 64, ... of the chunk; chunk sums are then added level by level."""
    big, one = 2.0**53, 1.0
    x = np.zeros(2048)
    x[0], x[1], x[64] = big, one, one
    y = np.ones(2048)
    # lanes = 1: ((big + 1) + ... + 1): each + 1 on 2^53 rounds to even -> big
    assert fxo.dot(x, y, 1) == big
    # 64 lanes: lane 0 holds big + 1 (entry 64) -> big, lane 1 holds 1; big + 1 -> big again
    assert fxo.dot(x, y, 64) == big
    x[64] = 0.0
    x[65] = one
    # lane 0 = big, lane 1 = 1 + 1 = 2 exactly, then big + 2 is exact; left to right it is big
    assert fxo.dot(x, y, 64) == big + 2.0
    assert fxo.dot(x, y, 1) == big
    # across chunks: chunk 0 = big, chunk 1 = 1, chunk 2 = 1: level 1 adds pairs (big + 1) -> big
    # and then big + 1 -> big; a plain running sum would give the same, a pairwise tree of
    # (1 + 1) first would give big + 2
    z = np.zeros(3 * 2048)
    z[0], z[2048], z[4096] = big, one, one
    w = np.ones(3 * 2048)
    assert fxo.dot(z, w, 64) == big


def test_oracle_cg_matches_dense_solve():
    model = meshes.column(8)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    K = np.column_stack([om.apply(op, e) for e in np.eye(om.neq)])
    b = om.body_load(op)
    x, r = om.cg(op, b, rtol=1e-13)
    assert r["converged"]
    xd = np.linalg.solve(K, b)
    assert np.max(np.abs(x - xd)) <= 1e-10 * np.max(np.abs(xd))
