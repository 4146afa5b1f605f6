"""Independent pins for three oracle functions (CPU only, no GPU):

* orc_symmetric_csr — against scipy.sparse: ((A + A^T) > 0) with the diagonal removed, as CSR;
* orc_tc_per_edge — both attributions against triangles enumerated by brute force from the
  undirected edge set, each triangle placed at the edge its form names (form 0: edge x->y of
  a triangle x<y<z in the (degree, id) order, the two vertices whose out-lists share z;
  form 1: edge x->z, the 1A2 attribution of ListIntersectionKernel, tc/bs_tc/bs_tc.cu:135-145);
* orc_tc_edge_model — closed forms on complete graphs, stars and friendship graphs, and
  hand-derived byte totals for K_3, K_4, K_5.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import load_oracle

orc = load_oracle()


# ---- orc_symmetric_csr ----------------------------------------------------------------------
def scipy_symmetric(pairs: np.ndarray, nodes: int):
    p = pairs.reshape(-1, 2).astype(np.int64)
    a = sp.coo_matrix((np.ones(len(p), dtype=np.int64), (p[:, 0], p[:, 1])), shape=(nodes, nodes)).tocsr()
    s = ((a + a.T) > 0).astype(np.int8).tolil()
    s.setdiag(0)
    s = s.tocsr()
    s.eliminate_zeros()
    s.sort_indices()
    return s.indptr.astype(np.int32), s.indices.astype(np.int32)


@pytest.mark.parametrize("seed,n,npairs,extra", [(1, 50, 400, 0), (2, 300, 5000, 7), (3, 2000, 30000, 100),
                                                 (4, 5, 40, 3)])
def test_symmetric_csr_matches_scipy(seed, n, npairs, extra):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, n, size=(npairs, 2)).astype(np.int32)
    # repeats, reversed repeats and self-loops on purpose
    p = np.concatenate([p, p[: npairs // 5], p[: npairs // 7, ::-1], np.repeat(np.arange(0, n, 3)[:, None], 2, 1)])
    rng.shuffle(p)
    nodes = n + extra  # trailing isolated vertices keep empty rows
    off, adj = orc.symmetric_csr(p.reshape(-1).astype(np.int32), nodes)
    e_off, e_adj = scipy_symmetric(p, nodes)
    assert np.array_equal(off, e_off)
    assert np.array_equal(adj, e_adj)


def test_symmetric_csr_self_loops_only_and_empty():
    off, adj = orc.symmetric_csr(np.array([2, 2, 0, 0], np.int32), 3)
    assert off.tolist() == [0, 0, 0, 0] and adj.size == 0
    off, adj = orc.symmetric_csr(np.zeros(0, np.int32), 2)
    assert off.tolist() == [0, 0, 0] and adj.size == 0


# ---- orc_tc_per_edge ------------------------------------------------------------------------
def graph_without_removals(n, p, rng):
    """Random simple graph on 0..n-1 in which every vertex has degree >= 2, so that the
    preprocessing keeps every id (relabel and removal are the identity)."""
    edges = {(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p}
    for v in range(n):  # a path through all vertices guarantees degree >= 2 inside 1..n-2
        if v + 1 < n:
            edges.add((v, v + 1))
    edges.add((0, n - 1))
    return sorted(edges)


def brute_force_per_edge(edges, n, off, adj):
    A = np.zeros((n, n), dtype=bool)
    deg = np.zeros(n, dtype=np.int64)
    for u, v in edges:
        A[u, v] = A[v, u] = True
        deg[u] += 1
        deg[v] += 1
    rank = lambda x: (deg[x], x)  # noqa: E731 — orientation order of tc/graph/graph.cpp:116-124
    form0 = np.zeros(adj.size, dtype=np.uint32)
    form1 = np.zeros(adj.size, dtype=np.uint32)

    def edge_index(a, b):
        row = adj[off[a]:off[a + 1]]
        k = int(np.searchsorted(row, b))
        assert k < row.size and row[k] == b, (a, b)
        return off[a] + k

    for t in itertools.combinations(range(n), 3):
        i, j, k = t
        if A[i, j] and A[j, k] and A[i, k]:
            x, y, z = sorted(t, key=rank)
            form0[edge_index(x, y)] += 1
            form1[edge_index(x, z)] += 1
    return form0, form1


@pytest.mark.parametrize("seed,n,p", [(1, 40, 0.3), (2, 120, 0.12), (3, 200, 0.06)])
def test_per_edge_both_forms_brute_force(seed, n, p):
    rng = np.random.default_rng(seed)
    edges = graph_without_removals(n, p, rng)
    pairs = np.asarray(edges, dtype=np.int32)[rng.permutation(len(edges))].reshape(-1)
    off, adj = orc.csr_from_pairs(pairs)
    assert off.size - 1 == n  # identity relabel, nothing removed
    b0, b1 = brute_force_per_edge(edges, n, off, adj)
    assert not np.array_equal(b0, b1)  # the two attributions differ here, so a swap is caught
    assert np.array_equal(orc.tc_per_edge(off, adj, 0), b0)
    assert np.array_equal(orc.tc_per_edge(off, adj, 1), b1)


def test_per_edge_k4_by_hand():
    # K_4, all degrees 3: orientation by id, 0->1,0->2,0->3,1->2,1->3,2->3 (CSR order).
    # Triangles {0,1,2},{0,1,3},{0,2,3},{1,2,3}: form 0 at x->y: 0->1 twice, 0->2, 1->2;
    # form 1 at x->z: 0->2, 0->3 twice, 1->3.
    off, adj = orc.csr_from_pairs(np.array(list(itertools.combinations(range(4), 2)), np.int32).reshape(-1))
    assert adj.tolist() == [1, 2, 3, 2, 3, 3]
    assert orc.tc_per_edge(off, adj, 0).tolist() == [2, 1, 0, 1, 0, 0]
    assert orc.tc_per_edge(off, adj, 1).tolist() == [0, 1, 2, 0, 1, 0]


# ---- orc_tc_edge_model ----------------------------------------------------------------------
def model_of(edges):
    off, adj = orc.csr_from_pairs(np.asarray(edges, np.int32).reshape(-1))
    return orc.tc_edge_model(off, adj), adj.size


@pytest.mark.parametrize("n", [3, 4, 5, 9, 30, 64])
def test_edge_model_complete_graph_closed_forms(n):
    # K_n orients by id: d+(i) = n-1-i. Sum over edges i->j of d+(j) counts the paths
    # i->j->k, i.e. C(n,3); sum of d+(i) is sum_i d+(i)^2 = (n-1)n(2n-1)/6.
    m, ne = model_of(list(itertools.combinations(range(n), 2)))
    assert ne == n * (n - 1) // 2
    assert m["wedges_v"] == n * (n - 1) * (n - 2) // 6
    assert m["wedges_u"] == (n - 1) * n * (2 * n - 1) // 6
    assert m["merge_bytes"] == 24 * ne + 4 * (m["wedges_u"] + m["wedges_v"])
    # never above the merge model, never below the endpoint/offset bytes plus the short lists
    assert m["alg_bytes"] <= m["merge_bytes"]


@pytest.mark.parametrize("n,expect", [(3, 84), (4, 188), (5, 352)])
def test_edge_model_small_complete_graphs_by_hand(n, expect):
    # Per edge 24 + 4*min(s + l, s*(1 + ceil(log2(l + 1)))), (s, l) the shorter and longer
    # out-lists. K_3 (d+ = 2,1,0): (2,1) -> 24+12, two edges with s = 0 -> 24 each: 84.
    # K_4 (3,2,1,0): (3,2) 44, (3,1) 36, (3,0) 24, (2,1) 36, (2,0) 24, (1,0) 24: 188.
    # K_5 (4,3,2,1,0): (4,3) 52, (4,2) 48, (4,1) 40, (4,0) 24, (3,2) 44, (3,1) 36, (3,0) 24,
    # (2,1) 36, (2,0) 24, (1,0) 24: 352.
    m, _ = model_of(list(itertools.combinations(range(n), 2)))
    assert m["alg_bytes"] == expect


def test_edge_model_star_is_empty():
    # every leaf has raw degree 1 and is removed, then the centre has no edge left
    m, ne = model_of([(0, i) for i in range(1, 50)])
    assert ne == 0 and m == {"wedges_v": 0, "wedges_u": 0, "alg_bytes": 0, "merge_bytes": 0}


@pytest.mark.parametrize("k", [1, 3, 40])
def test_edge_model_friendship_closed_forms(k):
    # A star whose leaves are paired by k blade edges: leaves (degree 2) point at the centre
    # (degree 2k >= 2), each blade lower id -> higher id. d+ = 2 (odd leaf), 1 (even leaf),
    # 0 (centre). Per blade: odd->even (2,1) 36 B, odd->centre and even->centre 24 B each.
    edges = []
    for i in range(k):
        a, b = 2 * i + 1, 2 * i + 2
        edges += [(0, a), (0, b), (a, b)]
    m, ne = model_of(edges)
    assert ne == 3 * k
    assert m["wedges_v"] == k
    assert m["wedges_u"] == 5 * k
    assert m["alg_bytes"] == 84 * k
    assert m["merge_bytes"] == 96 * k
