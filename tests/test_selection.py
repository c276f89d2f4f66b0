"""Hand-derived pins of contraction!'s integer control flow
(src/tensor_network.jl:152-155, :185-197, :248-316, :556-717).

Each network below is small enough that its schedule is worked out by hand in
the comments, step by step, from the reference text; the expected values are
those hand results, not the output of any implementation. They are checked
against the oracle (oracle/catn_oracle.py), the CPU emulator of the engine and,
marked gpu, the HIP library. Each pin is chosen so that a plausible slip fails
it:
  * select=1 (select_edge_min_dim_triangle, :262-288): strict `>` from -1,
    first candidate wins a tie (a `>=` picks the last one), and a later
    candidate with strictly more common neighbours beats the first;
  * dim_after_merge (:152-155): round to nearest, with dimensions 3 and 5 where
    floor, ceil and truncation all give a different first edge. Exact .5 ties
    cannot occur: the argument is log2(P_i P_j / d^2) for integer dimensions,
    which is an integer or irrational, never x.5, so ties-to-even and
    ties-away are indistinguishable on reachable inputs;
  * reverse (:608-619, `cld`): a 3-leg node i with idx_j_in_i = 1 is reversed
    (cld(3,2) = 2; fld would not), a 3-leg node j with idx_i_in_j = 1 is not
    (fld would), and a 2-leg node j is reversed whatever its index
    (cld(2,2) = 1);
  * duplicate bonds (:627-650): cross = idx_i_in_k > idx_j_in_k, both branches;
  * select=2 (:300-316): minimal i + j with strict `<`; among equal sums the
    first in node iteration order, which here is ascending id (the reference
    iterates a Julia Dict, whose order the build fixes this way);
  * Dmax-cut kept counts, in order, against the golden fixtures (test_engine).
All bonds have dimension 2 unless stated. Labels are given per tensor in leg
order; a label used once is an open leg (a negative sentinel neighbour).
"""
import numpy as np
import pytest

from oracle import catn_oracle as O


@pytest.fixture(scope="module", params=["oracle", "emu", pytest.param("gpu", marks=pytest.mark.gpu)])
def run(request, catn):
    """run(tensors, ixs, **kw) -> (edge_log, trace steps, trace dups, maxdim)."""
    if request.param == "oracle":
        def f(tensors, ixs, **kw):
            tn = O.tensor_network(tensors, ixs, **kw)
            s = O.Schedule()
            O.contraction(tn, s)
            return s.edges, s.steps, s.dups, tn.maxdim_intermediate
        return f
    ctx = request.getfixturevalue("emu_ctx" if request.param == "emu" else "gpu_ctx")

    def g(tensors, ixs, **kw):
        tn = catn.tensor_network(tensors, ixs, ctx=ctx, **kw)
        catn.contraction(tn)
        t = tn.plan_trace()
        return tn.edge_log(), t["steps"], t["dups"], tn.maxdim_intermediate
    return g


def _tensors(ixs, dims, seed=0):
    r = np.random.default_rng(seed)
    return [r.uniform(0.5, 1.5, [dims.get(l, 2) for l in ix]) for ix in ixs]


EXACT = dict(Dmax=-1, chi=10_000)


def test_select1_tie_first_candidate_wins(run):
    """Path 1-2-3-4: t1[a], t2[a,b], t3[b,c], t4[c]; neighbours 1:[2],
    2:[1,3], 3:[2,4], 4:[3].
    Costs (all legs dim 2: logdim = number of legs): (1,2) = 1+2-2 = 1,
    (2,3) = 2+2-2 = 2, (3,4) = 1. select_edge_init buckets, lexicographic:
    1: [(1,2), (3,4)], 2: [(2,3)].
    Step 1: both candidates have 0 common neighbours ((1,2): neigh(2)\\{1} =
    {3}, 1 is not a neighbour of 3; (3,4): neigh(4)\\{3} = {}); 0 > -1 takes
    (1,2), 0 > 0 fails for (3,4): (1,2). order(2) = 2 > order(1) = 1, so
    i = 2, j = 1; idx_j_in_i = 1 (1 < cld(2,2) = 1 fails: not reversed),
    idx_i_in_j = 1 >= cld(1,2) = 1: node 1 reversed. eat! leaves 2:[3].
    Step 2: bucket 1 = [(2,3), (3,4)] (re-added in sorted order; (2,3) now
    costs 1+2-2 = 1); again 0 and 0, first wins: (2,3). i = 3 (2 legs),
    j = 2: idx_j_in_i = find([2,4], 2) = 1, not reversed; node 2 (1 leg)
    reversed. eat! leaves 3:[4].
    Step 3: (3,4), i = 3, j = 4, j reversed."""
    ixs = [["a"], ["a", "b"], ["b", "c"], ["c"]]
    edges, steps, dups, _ = run(_tensors(ixs, {}), ixs, select=1, **EXACT)
    assert edges == [(1, 2), (2, 3), (3, 4)]
    assert steps == [(1, 2, 2, 1, 0, 1, 1, 1), (2, 3, 3, 2, 0, 1, 1, 1), (3, 4, 3, 4, 0, 1, 1, 1)]
    assert dups == []


def _graph_x(t4):
    # 1-2 and 2-3 pendant, 2-4, and the triangle 4-5-6
    return [["a"], ["a", "b", "c"], ["b"], t4, ["d", "f"], ["e", "f"]]


@pytest.mark.parametrize("t4,cross", [(["c", "d", "e"], 0), (["c", "e", "d"], 1)])
def test_select1_triangle_and_duplicate_cross(run, t4, cross):
    """t1[a], t2[a,b,c], t3[b], t4 (legs c,d,e or c,e,d), t5[d,f], t6[e,f]:
    bonds a:(1,2), b:(2,3), c:(2,4), d:(4,5), e:(4,6), f:(5,6).
    Costs deg(i)+deg(j)-2: (1,2) 2, (2,3) 2, (2,4) 4, (4,5) 3, (4,6) 3,
    (5,6) 2; bucket 2 = [(1,2), (2,3), (5,6)].
    Step 1, common neighbours: (1,2): neigh(2)\\{1} = {3,4}, neither has 1:
    0 (taken, 0 > -1); (2,3): {} : 0 (not > 0); (5,6): neigh(6)\\{5} = {4}, 4
    has 5: 1 > 0, taken. Edge (5,6) (select=0 would take (1,2)). Both have 2
    legs: i = 5, j = 6. idx_j_in_i = find([4,6], 6) = 2, 2 < 1 fails; node 6
    has 2 legs, idx_i_in_j = 2 >= cld(2,2) = 1: reversed to [5,4], so
    idx_i_in_j = 1.
    Re-pointing j's other neighbour k = 4 (l = 2): 4 is a neighbour of 5, a
    duplicate. With t4 = [c,d,e], node 4 = [2,5,6]: idx_i_in_k = 2, deleting 6
    gives idx_j_in_k = 3, cross = 2 > 3 = false. With t4 = [c,e,d], node 4 =
    [2,6,5]: idx_i_in_k = 3, idx_j_in_k = 2, cross = true."""
    ixs = _graph_x(t4)
    edges, steps, dups, _ = run(_tensors(ixs, {}), ixs, select=1, **EXACT)
    assert edges[0] == (5, 6)
    assert steps[0] == (5, 6, 5, 6, 0, 1, 2, 1)
    assert dups[0] == (0, 4, cross)


def test_dim_after_merge_rounds_to_nearest(run):
    """t1[a], t2[a,b], t3[b] with dim(a) = 3, dim(b) = 5.
    dim_after_merge(1,2) = log2 3 + (log2 3 + log2 5) - 2 log2 3 = log2 5 =
    2.32 -> 2; dim_after_merge(2,3) = (log2 3 + log2 5) + log2 5 - 2 log2 5 =
    log2 3 = 1.58 -> 2. Both in bucket 2, inserted (1,2) then (2,3): select=0
    takes (1,2). (floor or truncation: (2,3) costs 1 and goes first; ceil:
    (1,2) costs 3 and (2,3) goes first.) Node 2 eats node 1, keeping one leg
    of dimension 5: maxdim_intermediate = max(-1, 2, 5) = 5; then (2,3) costs
    log2 5 + log2 5 - 2 log2 5 = 0."""
    ixs = [["a"], ["a", "b"], ["b"]]
    edges, steps, _, maxdim = run(_tensors(ixs, {"a": 3, "b": 5}), ixs, select=0, **EXACT)
    assert edges == [(1, 2), (2, 3)]
    assert steps[0][:4] == (1, 2, 2, 1)
    assert maxdim == 5


def test_reverse_rule_uses_cld(run):
    """t1[a,x,y], t2[a,z,w] (x, y, z, w open): one bond, (1,2), cost
    3 + 3 - 2 = 4. Equal orders: i = 1, j = 2; idx_j_in_i = idx_i_in_j = 1.
    Node i: 1 < cld(3,2) = 2, reversed (legs [2,-1,-2] become [-2,-1,2], so
    idx_j_in_i = 3). Node j: 1 >= cld(3,2) = 2 fails, not reversed. With fld
    both decisions flip."""
    ixs = [["a", "x", "y"], ["a", "z", "w"]]
    _, steps, _, _ = run(_tensors(ixs, {}), ixs, select=0, **EXACT)
    assert steps == [(1, 2, 1, 2, 1, 0, 3, 1)]


def test_reverse_rule_two_leg_node_j(run):
    """t1[a,x,y], t2[a,z]: i = 1 (3 legs), j = 2 (2 legs). Node i reversed
    as above (idx_j_in_i -> 3); node j: idx_i_in_j = 1 >= cld(2,2) = 1,
    reversed to [-3, 1], so idx_i_in_j = 2."""
    ixs = [["a", "x", "y"], ["a", "z"]]
    _, steps, _, _ = run(_tensors(ixs, {}), ixs, select=0, **EXACT)
    assert steps == [(1, 2, 1, 2, 1, 1, 3, 2)]


def test_select2_min_index_sum(run):
    """Bonds (1,4), (2,3), (3,4), (1,2): t1[a,d], t2[d,b], t3[b,c], t4[a,c]
    (a:(1,4), b:(2,3), c:(3,4), d:(1,2)). Sums: 5, 5, 7, 3: the first step
    takes (1,2). i = 1 (equal orders), j = 2; node 1 eats node 2 and, 3
    being no neighbour of 1, bond b is re-pointed: 1 = [4, 3] after eat!
    (node 1's legs [4,2] with 2 removed, then node 2's other neighbour 3).
    Step 2: bonds (1,4), (1,3), (3,4); sums 5, 4, 7: (1,3). Step 3: 1 and 4
    now share two bonds, merged into one: (1,4)."""
    ixs = [["a", "d"], ["d", "b"], ["b", "c"], ["a", "c"]]
    edges, _, _, _ = run(_tensors(ixs, {}), ixs, select=2, **EXACT)
    assert edges == [(1, 2), (1, 3), (1, 4)]


def test_select2_ties_take_the_lowest_node_id(run):
    """Bonds a:(1,4), b:(2,3) only (two components): sums 5 and 5. Node
    iteration in ascending id meets (1,4) first and strict `<` keeps it."""
    ixs = [["a"], ["b"], ["b"], ["a"]]
    edges, _, _, _ = run(_tensors(ixs, {}), ixs, select=2, **EXACT)
    assert edges == [(1, 4), (2, 3)]


def test_planted_mistakes_fail_the_pins(monkeypatch):
    """The pins catch plausible slips in the oracle: `>=` in the triangle
    selector, floor in dim_after_merge, fld in the reverse rule."""
    ixs = [["a"], ["a", "b"], ["b", "c"], ["c"]]

    def ge_triangle(tn):
        cost = min(k for k, v in tn.edge_count.items() if v)
        tn.maxdim_intermediate = max(tn.maxdim_intermediate, cost)
        best, bestc = tn.edge_count[cost][0], -1
        for p in tn.edge_count[cost]:
            i, j = p
            neigh_j = {nb for nb in tn.tensors[j].neighbor if nb > 0 and nb != i}
            both = sum(1 for k in neigh_j if O.find_neighbor(tn.tensors[k], i) > 0)
            if both >= bestc:
                bestc, best = both, p
        return best[0], best[1]
    with monkeypatch.context() as m:
        m.setattr(O, "select_edge_min_dim_triangle", ge_triangle)
        s = O.Schedule()
        O.contraction(O.tensor_network(_tensors(ixs, {}), ixs, select=1, **EXACT), s)
        assert s.edges != [(1, 2), (2, 3), (3, 4)]

    ixr = [["a"], ["a", "b"], ["b"]]
    with monkeypatch.context() as m:
        def floor_dim(tn, i, j):
            ni = tn.tensors[i]
            idx = O.find_neighbor(ni, j)
            return int(np.floor(O.logdim(ni) + O.logdim(tn.tensors[j]) - 2 * O.logdim(ni, idx)))
        m.setattr(O, "dim_after_merge", floor_dim)
        s = O.Schedule()
        O.contraction(O.tensor_network(_tensors(ixr, {"a": 3, "b": 5}), ixr, select=0, **EXACT), s)
        assert s.edges[0] != (1, 2)

    ixc = [["a", "x", "y"], ["a", "z", "w"]]
    with monkeypatch.context() as m:
        m.setattr(O, "_cld", lambda a, b: a // b)
        s = O.Schedule()
        O.contraction(O.tensor_network(_tensors(ixc, {}), ixc, select=0, **EXACT), s)
        assert s.steps != [(1, 2, 1, 2, 1, 0, 3, 1)]
