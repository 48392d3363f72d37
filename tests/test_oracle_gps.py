"""GPS node relabelling (core/meshRelabler.cpp:57-530): pins of the oracle, and the native
relabeller (od_relabel_gps, host code) against it.

Pins of the oracle:
- a 6-node chain of 3 tets, traced by hand through the reference's steps (pseudo-peripheral
  pair 0 and 5, vertices 0, 2, 3, 5 fixed by both level structures, one component {4, 1}
  placed on the end structure's levels, selectSecond set, no interchange): labels 0..5;
  and the chain with a branch tet, whose second component takes the start structure's
  levels (the unequal-count branch) and whose restored arcs hold duplicates;
- shuffled tet chains: every edge joins nodes at most 5 labels apart (levels of width 3);
- box meshes with shuffled node ids and scipy Delaunay meshes of random points: the labels
  are a permutation; the nodes labelled first and last both have eccentricity equal to the
  graph diameter (scipy BFS), which the pseudo-peripheral search must reach on these graphs;
  the bandwidth is below twice the widest BFS level from those nodes and far below the
  shuffled input's;
- the elements come back relabelled and sorted by label sum, stably;
- a disconnected mesh is rejected.
Exact labels of the reference binary depend on its standard library's unordered_set order
and unstable sorts; the restatement fixes those orders (see the oracle), so they are not
pinned against the reference itself.
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path
from scipy.spatial import Delaunay

import oracle as O
from opendeformer_amd import mesh as M


def chain(n):
    return np.array([[i, i + 1, i + 2, i + 3] for i in range(n - 3)], np.int32)


def graph(n, tets):
    r = np.concatenate([tets[:, a] for a in range(4) for b in range(4) if a != b])
    c = np.concatenate([tets[:, b] for a in range(4) for b in range(4) if a != b])
    return csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))


def bandwidth(lab, tets):
    t = lab[tets]
    return int((t.max(1) - t.min(1)).max())


def delaunay_mesh(n, seed):
    pts = np.random.default_rng(seed).random((n, 3))
    return Delaunay(pts).simplices.astype(np.int32)


def cases():
    rng = np.random.default_rng(7)
    out = []
    for dims in [(6, 3, 2), (10, 4, 4), (5, 5, 5), (12, 2, 3)]:
        m = M.box_mesh(*dims, 1.0, 1.0, 1.0, seed=1)
        out.append((f"box{dims}", m.n_nodes, m.tets))
        perm = rng.permutation(m.n_nodes)
        out.append((f"box{dims}-shuffled", m.n_nodes, perm[m.tets].astype(np.int32)))
    for seed in (1, 2):
        out.append((f"delaunay{seed}", 300, delaunay_mesh(300, seed)))
    return out


def test_hand_traced_chain(oracle):
    lab, el = O.gps(6, chain(6))
    assert lab.tolist() == [0, 1, 2, 3, 4, 5]
    assert el.tolist() == chain(6).tolist()


def test_hand_traced_branch(oracle):
    """The chain with a tet (2, 3, 6, 7) hung on its middle, traced by hand: pseudo-peripheral
    pair 0 and 7; vertices 7, 3, 2, 0 fixed by both structures; components {6} (equal counts:
    end levels, selectSecond) and {5, 4, 1} (counts 2 against 3: start levels); the restored
    arcs give degrees 3 4 8 9 4 3 3 5; numbering 0 | 1, 3, 2, 6 | 4, 5, 7."""
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4], [2, 3, 4, 5], [2, 3, 6, 7]], np.int32)
    lab, el = O.gps(8, tets)
    assert lab.tolist() == [0, 1, 3, 2, 5, 6, 4, 7]
    assert el.tolist() == [[0, 1, 3, 2], [1, 3, 2, 5], [3, 2, 5, 6], [3, 2, 4, 7]]


def test_shuffled_chains_stay_banded(oracle):
    rng = np.random.default_rng(3)
    for _ in range(100):
        n = int(rng.integers(8, 40))
        perm = rng.permutation(n)
        tets = perm[chain(n)]
        lab, _ = O.gps(n, tets)
        assert sorted(lab.tolist()) == list(range(n))
        assert bandwidth(lab, tets) <= 5
        pos = np.argsort(perm)                  # chain position of each id
        assert pos[np.argsort(lab)[0]] in (0, n - 1)


@pytest.mark.parametrize("name,n,tets", cases(), ids=[c[0] for c in cases()])
def test_level_structure_properties(oracle, name, n, tets):
    lab, el = O.gps(n, tets)
    assert sorted(lab.tolist()) == list(range(n))
    D = shortest_path(graph(n, tets), unweighted=True)
    diam = D.max()
    first, last = int(np.argmin(lab)), int(np.argmax(lab))
    assert D[first].max() == diam and D[last].max() == diam
    width = max(np.bincount(D[first].astype(int)).max(), np.bincount(D[last].astype(int)).max())
    assert bandwidth(lab, tets) <= 2 * width - 1
    if "shuffled" in name:
        assert bandwidth(lab, tets) < bandwidth(np.arange(n), tets) / 2
    # elements: relabelled, sorted by label sum, stable
    rel = lab[tets]
    order = np.argsort(rel.sum(1), kind="stable")
    assert np.array_equal(el, rel[order])


def test_disconnected_mesh_is_rejected(oracle):
    tets = np.array([[0, 1, 2, 3], [4, 5, 6, 7]], np.int32)
    with pytest.raises(ValueError):
        O.gps(8, tets)


def native_cases():
    branch = np.array([[0, 1, 2, 3], [1, 2, 3, 4], [2, 3, 4, 5], [2, 3, 6, 7]], np.int32)
    perm = np.random.default_rng(9).permutation(40)
    return cases() + [("chain6", 6, chain(6)), ("branch", 8, branch), ("chain40-shuffled", 40, perm[chain(40)])]


@pytest.mark.parametrize("name,n,tets", native_cases(), ids=[c[0] for c in native_cases()])
def test_native_relabeller_equals_oracle(oracle, name, n, tets):
    lab, el = O.gps(n, tets)
    X = np.zeros((n, 3))
    m = M.TetMesh(X=X, tets=tets.copy(), facets=np.zeros((0, 3), np.int32), shape=(0, 0, 0), size=(0, 0, 0))
    g = M.gps_relabel(m)
    assert np.array_equal(g.extra["perm"], lab)
    assert np.array_equal(g.tets, el)


def test_native_relabeller_rejects_a_disconnected_mesh():
    tets = np.array([[0, 1, 2, 3], [4, 5, 6, 7]], np.int32)
    m = M.TetMesh(X=np.zeros((8, 3)), tets=tets, facets=np.zeros((0, 3), np.int32), shape=(0, 0, 0), size=(0, 0, 0))
    with pytest.raises(ValueError, match="not connected"):
        M.gps_relabel(m)
