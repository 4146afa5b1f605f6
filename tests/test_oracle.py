"""Pins the CPU oracle (CPU only, no GPU): closed forms, brute force, independent libraries
(networkx, scipy, numpy), the reference's own known-answer invariants, and hand-derived
fixtures for the preprocessing semantics of tc/graph/graph.cpp."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import load_oracle

orc = load_oracle()


# ---- helpers -------------------------------------------------------------------------------
def pairs_of(edges):
    return np.asarray(edges, dtype=np.int32).reshape(-1)


def count_all_forms(pairs):
    off, adj = orc.csr_from_pairs(pairs)
    a = orc.tc_count_1a2(off, adj)
    b = orc.tc_count_intersect(off, adj)
    c = orc.tc_count_2a1(off, adj)
    assert a == b == c, (a, b, c)
    return a


def brute_force(edges, n):
    A = np.zeros((n, n), dtype=bool)
    for u, v in edges:
        if u != v:
            A[u, v] = A[v, u] = True
    t = 0
    for i, j, k in itertools.combinations(range(n), 3):
        t += A[i, j] and A[j, k] and A[i, k]
    return t


def random_graph(n, p, rng):
    return [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]


# ---- closed forms ----------------------------------------------------------------------------
@pytest.mark.parametrize("n", [3, 4, 5, 8, 20])
def test_complete_graph(n):
    edges = list(itertools.combinations(range(n), 2))
    assert count_all_forms(pairs_of(edges)) == n * (n - 1) * (n - 2) // 6


@pytest.mark.parametrize("k", [4, 5, 9, 30])
def test_wheel(k):
    edges = [(0, i) for i in range(1, k + 1)] + [(i, i % k + 1) for i in range(1, k + 1)]
    assert count_all_forms(pairs_of(edges)) == k


@pytest.mark.parametrize("k", [1, 2, 7])
def test_friendship(k):
    edges = []
    for i in range(k):
        a, b = 2 * i + 1, 2 * i + 2
        edges += [(0, a), (0, b), (a, b)]
    assert count_all_forms(pairs_of(edges)) == k


@pytest.mark.parametrize("a,b", [(1, 5), (3, 4), (10, 10)])
def test_complete_bipartite_is_triangle_free(a, b):
    edges = [(i, a + j) for i in range(a) for j in range(b)]
    assert count_all_forms(pairs_of(edges)) == 0


def test_tree_is_triangle_free():
    rng = np.random.default_rng(3)
    edges = [(i, int(rng.integers(0, i))) for i in range(1, 500)]
    assert count_all_forms(pairs_of(edges)) == 0


def test_single_triangle():
    assert count_all_forms(pairs_of([(0, 1), (1, 2), (2, 0)])) == 1


# ---- brute force and independent libraries --------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_brute_force_random(seed):
    rng = np.random.default_rng(seed)
    n = 40
    edges = random_graph(n, 0.25, rng)
    assert count_all_forms(pairs_of(edges)) == brute_force(edges, n)


def test_networkx_and_scipy_agree():
    nx = pytest.importorskip("networkx")
    sp = pytest.importorskip("scipy.sparse")
    G = nx.gnm_random_graph(600, 9000, seed=5)
    edges = list(G.edges())
    ours = count_all_forms(pairs_of(edges))
    assert ours == sum(nx.triangles(G).values()) // 3
    A = nx.to_scipy_sparse_array(G, format="csr").astype(np.int64)
    assert ours == int((A @ A).multiply(A).sum()) // 6


def test_count_invariant_under_relabel_reversal_duplicates():
    rng = np.random.default_rng(11)
    edges = random_graph(120, 0.1, rng)
    base = count_all_forms(pairs_of(edges))
    perm = rng.permutation(10_000)[:120] + 1  # sparse, 1-based ids
    e2 = [(int(perm[v]), int(perm[u])) for u, v in edges]  # relabelled and reversed
    e2 += e2[:50] + [(b, a) for a, b in e2[50:90]] + [(int(perm[3]), int(perm[3]))]
    rng.shuffle(e2)
    assert count_all_forms(pairs_of(e2)) == base


# ---- preprocessing semantics (tc/graph/graph.cpp) -------------------------------------------
def test_preprocess_hand_fixture():
    # 1-based ids with a gap (6 absent), a reversed duplicate, a repeated pair and a self-loop.
    # Derivation: relabel 1..5,7 -> 0..5; raw degrees [4,4,6,2,1,1]; vertices 4,5 removed;
    # orientation by (degree, id); sorted, deduplicated: 0->1, 0->2, 1->2, 3->2.
    edges = [(1, 2), (2, 3), (3, 1), (3, 4), (4, 5), (7, 3), (2, 1), (3, 3), (1, 2)]
    off, adj = orc.csr_from_pairs(pairs_of(edges))
    assert off.tolist() == [0, 2, 3, 3, 4]
    assert adj.tolist() == [1, 2, 2, 2]
    assert orc.tc_count_intersect(off, adj) == 1


def test_preprocess_zero_based_self_loop_only():
    off, adj = orc.csr_from_pairs(pairs_of([(0, 0), (0, 1)]))
    assert off.tolist() == [0, 0] and adj.size == 0


def test_preprocess_empty():
    off, adj = orc.csr_from_pairs(np.zeros(0, dtype=np.int32))
    assert off.tolist() == [0] and adj.size == 0


def test_relabel_is_dense_rank():
    rng = np.random.default_rng(2)
    for base in (0, 1):
        p = rng.integers(base, 5000, size=4000).astype(np.int32)
        if base == 1:
            p[p == 0] = 1
        out, nodes = orc.relabel(p)
        uniq = np.unique(p)
        assert nodes == uniq.size
        assert np.array_equal(out, np.searchsorted(uniq, p).astype(np.int32))


def test_orientation_properties():
    rng = np.random.default_rng(7)
    n = 300
    raw = rng.integers(0, n, size=(4000, 2)).astype(np.int32)
    p, nodes = orc.relabel(raw.reshape(-1))
    deg = np.bincount(p, minlength=nodes)
    keep = np.flatnonzero(deg > 1)
    newid = -np.ones(nodes, dtype=np.int64)
    newid[keep] = np.arange(keep.size)
    off, adj = orc.tc_preprocess(p, nodes)
    assert off.size - 1 == keep.size
    rows = np.repeat(np.arange(off.size - 1), np.diff(off))
    # rows strictly increasing, no self-loops
    for u in range(off.size - 1):
        r = adj[off[u]:off[u + 1]]
        assert np.all(np.diff(r) > 0)
    assert np.all(rows != adj)
    # orientation: (raw degree, id) increases along every edge
    old = keep  # new id -> relabelled id
    du, dv = deg[old[rows]], deg[old[adj]]
    assert np.all((du < dv) | ((du == dv) & (old[rows] < old[adj])))
    # every simple edge between kept vertices appears exactly once
    pp = p.reshape(-1, 2)
    ok = (deg[pp[:, 0]] > 1) & (deg[pp[:, 1]] > 1) & (pp[:, 0] != pp[:, 1])
    und = {tuple(sorted((int(newid[a]), int(newid[b])))) for a, b in pp[ok]}
    got = {tuple(sorted((int(a), int(b)))) for a, b in zip(rows, adj)}
    assert und == got


def test_per_edge_forms_sum_to_count():
    rng = np.random.default_rng(4)
    edges = random_graph(200, 0.08, rng)
    off, adj = orc.csr_from_pairs(pairs_of(edges))
    t = orc.tc_count_intersect(off, adj)
    assert int(orc.tc_per_edge(off, adj, 0).sum()) == t
    assert int(orc.tc_per_edge(off, adj, 1).sum()) == t


def test_edge_range_partition():
    rng = np.random.default_rng(9)
    edges = random_graph(250, 0.1, rng)
    off, adj = orc.csr_from_pairs(pairs_of(edges))
    m = adj.size
    t = orc.tc_count_1a2(off, adj)
    cuts = [0, m // 3, m // 2, m]
    assert sum(orc.tc_count_1a2(off, adj, a, b) for a, b in zip(cuts, cuts[1:])) == t


# ---- set intersection ---------------------------------------------------------------------
def reference_style_dataset(num_a, num_b, len_a, len_b, len_c, rng):
    """The generator invariant of set_intersection/data_generator/data.cpp:21-82: every A set
    equals `tmp`, every B set holds the len_c smallest... here: len_c elements of tmp plus
    fresh values outside tmp; every A x B intersection is exactly len_c."""
    tmp = np.sort(rng.choice(10 * (len_a + len_b), size=len_a, replace=False)).astype(np.uint32)
    sets = [tmp] * num_a
    pool = np.setdiff1d(np.arange(10 * (len_a + len_b), dtype=np.uint32), tmp)
    for _ in range(num_b):
        fresh = rng.choice(pool, size=len_b - len_c, replace=False)
        sets.append(np.sort(np.concatenate([tmp[:len_c], fresh])).astype(np.uint32))
    offs = np.zeros(len(sets) + 1, dtype=np.uint32)
    np.cumsum([s.size for s in sets], out=offs[1:])
    return offs, np.concatenate(sets).astype(np.uint32)


@pytest.mark.parametrize("len_c", [0, 1, 17])
def test_si_tile_reference_invariant(len_c):
    rng = np.random.default_rng(len_c)
    offs, el = reference_style_dataset(3, 5, 40, 25, len_c, rng)
    out = orc.si_tile(offs, el, 3, 5)
    assert np.all(out == len_c)


def test_si_pairs_bruteforce():
    rng = np.random.default_rng(1)
    lists = [np.unique(rng.integers(0, 300, size=rng.integers(0, 60))).astype(np.uint32) for _ in range(80)]
    offs = np.zeros(81, dtype=np.uint32)
    np.cumsum([x.size for x in lists], out=offs[1:])
    el = np.concatenate(lists).astype(np.uint32)
    pairs = rng.integers(0, 80, size=(500, 2)).astype(np.uint32)
    out = orc.si_pairs(offs, el, pairs)
    exp = [len(set(lists[a].tolist()) & set(lists[b].tolist())) for a, b in pairs]
    assert out.tolist() == exp


# ---- sorted lookup ------------------------------------------------------------------------
@pytest.mark.parametrize("n", [0, 1, 2, 7, 1000])
def test_lookup_first_occurrence(n):
    rng = np.random.default_rng(n)
    data = np.sort(rng.integers(0, max(2 * n, 4), size=n)).astype(np.uint32)  # duplicates
    keys = rng.integers(0, max(2 * n, 4) + 2, size=3000).astype(np.uint32)
    got = orc.lookup(data, keys)
    idx = np.searchsorted(data, keys, side="left")
    hit = (idx < n) & (data[np.minimum(idx, max(n - 1, 0))] == keys) if n else np.zeros(keys.size, bool)
    exp = np.where(hit, idx, -1)
    assert np.array_equal(got, exp)


def test_lookup_signed():
    data = np.array([-5, -5, -1, 0, 0, 3, 2**31 - 1], dtype=np.int32)
    keys = np.array([-5, -4, 0, 3, 2**31 - 1, -2**31], dtype=np.int32)
    assert orc.lookup(data, keys, signed=True).tolist() == [0, -1, 3, 5, 6, -1]
