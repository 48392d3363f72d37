"""Pins the oracle's pattern, 78-entry element map and SpMDV without the reference.

The canonical layout (core/sparseMatrix.h:90-274) is re-derived here by brute
force from a dense boolean adjacency and the stated layout rules; the element
map (material/fullOrderNonlinear.h:73-110) is checked against it entry by
entry, and SpMDV (core/numerMethod.h:134-262) against a dense product.
"""
import numpy as np
import pytest

from opendeformer_amd import mesh as M
from opendeformer_amd import workloads as W


def global_index(node, axis, con):
    idx = 3 * node + axis
    k = np.searchsorted(con, idx)
    if k < len(con) and con[k] == idx:
        return -1
    return idx - k


def brute_layout(ndof, tets, con):
    """Dense lower pattern -> (blockPcol, blockColumnOris, blockRows, entry-position dict)."""
    P = np.zeros((ndof, ndof), bool)
    for t in tets:
        g = [global_index(n, a, con) for n in t for a in range(3)]
        g = [x for x in g if x >= 0]
        for r in g:
            for c in g:
                if r >= c:
                    P[r, c] = True
    nbc = ndof // 3
    pcol, oris, rows, pos = [0], [0], [], {}
    off = 0
    for J in range(nbc):
        c0 = 3 * J
        cols = P[:, c0:c0 + 3]
        starts = []
        if cols[c0:c0 + 3].any():
            starts.append(c0)
        r = c0 + 3
        while r < ndof:
            if cols[r:r + 3].any():
                starts.append(r)
            r += 3
        for s in starts:
            rows.append(s)
            if s == c0:
                for i in range(3):           # packed (0,0),(1,0),(2,0),(1,1),(2,1),(2,2)
                    for j in range(i, 3):
                        pos[(c0 + j, c0 + i)] = off
                        off += 1
            else:
                w = min(3, ndof - s)
                for i in range(3):           # column-major inside the block
                    for j in range(w):
                        pos[(s + j, c0 + i)] = off
                        off += 1
        pcol.append(len(rows)); oris.append(off)
    for c in range(3 * nbc, ndof):           # remained scalar columns
        for r in range(c, ndof):
            if P[r, c]:
                rows.append(r); pos[(r, c)] = off; off += 1
        pcol.append(len(rows)); oris.append(off)
    return np.array(pcol), np.array(oris), np.array(rows), pos, P


def brute_map(tets, con, pos):
    out = np.full((len(tets), 78), -1, np.int64)
    for e, t in enumerate(tets):
        g = [global_index(n, a, con) for n in t for a in range(3)]
        for c in range(12):
            for r in range(c, 12):
                li = c * 12 - ((c - 1) * c) // 2 + (r - c)
                if g[c] >= 0 and g[r] >= 0:
                    out[e, li] = pos[(max(g[r], g[c]), min(g[r], g[c]))]
    return out


CASES = {
    "box_small": lambda: (W.box_small().X, W.box_small().tets, W.box_small().fixed_dofs),
    # partial constraints exercise the truncated last block and the remained scalar columns
    "partial": lambda: (M.box_mesh(3, 2, 2, 1, 1, 1, seed=5).X, M.box_mesh(3, 2, 2, 1, 1, 1, seed=5).tets,
                        np.array([0, 1, 2, 4, 7, 8, 13], np.int32)),
    "color8": lambda: (W.box_small("color8").X, W.box_small("color8").tets, W.box_small("color8").fixed_dofs),
    "color5": lambda: (W.box_small("color5").X, W.box_small("color5").tets, W.box_small("color5").fixed_dofs),
}


@pytest.mark.parametrize("colors", [8, 5, 4])
def test_multicolour_orderings_are_proper(colors):
    """No tet edge joins two nodes of one colour, so the IC(0) dependency levels of a
    colour-grouped numbering number at most the colours."""
    m = M.box_mesh(9, 6, 5, 1.0, 0.6, 0.5, seed=3)
    perm = M.multicolor_permutation(m, colors)
    n = np.arange(m.n_nodes)
    k, j, i = n % 6, (n // 6) % 7, n // 42
    color = {8: (i + 2 * j + 4 * k) % 8, 5: (i + j + 2 * k) % 5, 4: (i + j + k) % 4}[colors]
    assert len(np.unique(color)) == colors
    for a in range(4):
        for b in range(a + 1, 4):
            assert (color[m.tets[:, a]] != color[m.tets[:, b]]).all()
    # the relabelling groups colours: new ids of colour c precede those of colour c + 1
    new_color = np.empty_like(color); new_color[perm] = color
    assert (np.diff(new_color) >= 0).all()
    with pytest.raises(ValueError):          # a tet is a 4-clique: no 3-colouring exists
        M.multicolor_permutation(m, 3)


@pytest.mark.parametrize("case", sorted(CASES))
def test_pattern_and_map_match_brute_force(oracle, case):
    X, tets, con = CASES[case]()
    s = oracle.System(X, tets, con)
    pcol, oris, rows, emap = s.pattern()
    bp, bo, br, pos, P = brute_layout(s.ndof, tets, np.asarray(con))
    assert np.array_equal(pcol, bp)
    assert np.array_equal(oris, bo)
    assert np.array_equal(rows, br)
    assert np.array_equal(emap, brute_map(tets, np.asarray(con), pos))
    assert s.n_values == len(pos)
    if case != "partial":
        # with whole-node constraints every stored entry is hit by some element pair
        hit = np.zeros(s.n_values, bool)
        hit[emap[emap >= 0]] = True
        assert hit.all()


def test_box_counts_match_workload_sizes(oracle):
    for w, nodes, tets, dof, blocks, vals in [(W.box_small(), 525, 1920, 1500, 3159, 26931)]:
        s = oracle.System(w.X, w.tets, w.fixed_dofs)
        assert (w.n_nodes, w.n_tets, s.ndof, s.n_blocks, s.n_values) == (nodes, tets, dof, blocks, vals)
    m = M.box_mesh(80, 16, 16, 2.0, 0.4, 0.4, seed=2)
    assert (m.n_nodes, m.n_tets) == (23409, 122880)
    assert (M.tet_signed_volumes(m.X, m.tets) > 0).all()


def test_constrainer_plane_selects_x0_face(oracle):
    w = W.box_small()
    con = oracle.constrain_plane(w.X, w.facets, [0, 0, 0], [1, 0, 0])
    assert np.array_equal(con, w.fixed_dofs)
    assert len(con) == 75


def test_spmdv_matches_dense_and_accumulates(oracle):
    w = W.box_small()
    s = oracle.System(w.X, w.tets, w.fixed_dofs)
    rng = np.random.default_rng(3)
    vals = rng.standard_normal(s.n_values)
    _, _, _, pos, _ = brute_layout(s.ndof, w.tets, w.fixed_dofs)
    A = np.zeros((s.ndof, s.ndof))
    for (r, c), p in pos.items():
        A[r, c] = A[c, r] = vals[p]
    x = rng.standard_normal(s.ndof)
    y0 = rng.standard_normal(s.ndof)
    y = s.spmdv(vals, x, y0.copy())
    np.testing.assert_allclose(y, y0 + A @ x, rtol=1e-12, atol=1e-12 * np.abs(A @ x).max())


def test_spmdv_partial_constraints_dense(oracle):
    X, tets, con = CASES["partial"]()
    s = oracle.System(X, tets, con)
    rng = np.random.default_rng(4)
    vals = rng.standard_normal(s.n_values)
    _, _, _, pos, _ = brute_layout(s.ndof, tets, con)
    A = np.zeros((s.ndof, s.ndof))
    for (r, c), p in pos.items():
        A[r, c] = A[c, r] = vals[p]
    x = rng.standard_normal(s.ndof)
    np.testing.assert_allclose(s.spmdv(vals, x), A @ x, rtol=1e-12, atol=1e-12)
    d = s.diagonal(vals)
    np.testing.assert_array_equal(d, np.diag(A))
