"""Seeded box-grid tetrahedral meshes (Kuhn 6-tet split) and node relabelling.

The reference ships no mesher for box grids (its Delaunay mesher lives in
mesher/delMesher.cpp and consumes surface meshes), so the workloads here are
generated directly:

* node (i, j, k), 0 <= i <= nx etc., X = (i*Lx/nx, j*Ly/ny, k*Lz/nz);
* node number n = k + (nz+1) * (j + (ny+1) * i)  (x slowest, so the x = 0 face
  is the first (ny+1)(nz+1) nodes);
* interior nodes (0<i<nx, 0<j<ny, 0<k<nz) get an independent U(-0.1, 0.1)
  times the cell size jitter per axis from a counter-based splitmix64 stream
  keyed by (seed, node, axis); boundary nodes stay exact so the reference's
  plane constraint test |(v - o) . n| < 1e-4 (core/constrainer.cpp:57-72)
  selects exactly the x = 0 face;
* every cube is split along its v000 -> v111 diagonal into 6 tets, one per
  axis permutation, each stored with positive orientation;
* boundary facets: two outward triangles per boundary square, split along the
  same diagonal as the adjacent tet faces (needed by Constrainer, which scans
  facets only).

Everything is plain numpy; the arrays feed both the oracle and the HIP library.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """Vectorised splitmix64 finaliser over uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, idx) -> np.ndarray:
    """Counter-based U[0,1) doubles: one value per index, keyed by (seed, stream)."""
    idx = np.asarray(idx, dtype=np.uint64)
    key = _splitmix64(np.array([(seed & 0xFFFFFFFF) | ((stream & 0xFFFF) << 40)], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        h = _splitmix64((idx * np.uint64(0x2545F4914F6CDD1D) + key) & _M64)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# stream ids used by the workload generators
STREAM_JITTER = 1
STREAM_YOUNG = 2
STREAM_IMPULSE = 3
STREAM_INSTANCE = 4


@dataclass
class TetMesh:
    X: np.ndarray            # (nNodes, 3) float64 rest positions
    tets: np.ndarray         # (nTets, 4) int32
    facets: np.ndarray       # (nFacets, 3) int32 boundary triangles
    shape: tuple = ()        # (nx, ny, nz) for grid meshes
    size: tuple = ()         # (Lx, Ly, Lz)
    extra: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def n_tets(self) -> int:
        return int(self.tets.shape[0])


def box_mesh(nx: int, ny: int, nz: int, Lx: float, Ly: float, Lz: float,
             seed: int = 0, jitter: float = 0.1) -> TetMesh:
    """Kuhn-split box grid with seeded interior jitter (see module docstring)."""
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I = I.ravel(); J = J.ravel(); K = K.ravel()
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    X = np.stack([I * hx, J * hy, K * hz], axis=1).astype(np.float64)
    nn = X.shape[0]
    if jitter > 0:
        interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        ids = np.nonzero(interior)[0].astype(np.uint64)
        cell = np.array([hx, hy, hz])
        for ax in range(3):
            u = uniform01(seed, STREAM_JITTER, ids * np.uint64(3) + np.uint64(ax))
            X[ids, ax] += (2.0 * u - 1.0) * jitter * cell[ax]

    def nid(i, j, k):
        return k + (nz + 1) * (j + (ny + 1) * i)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.ravel(); cj = cj.ravel(); ck = ck.ravel()
    v000 = nid(ci, cj, ck)
    e = [np.array([1, 0, 0]), np.array([0, 1, 0]), np.array([0, 0, 1])]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    parity = [1, -1, -1, 1, 1, -1]
    cols = []
    for p, sgn in zip(perms, parity):
        a = e[p[0]]
        b = e[p[0]] + e[p[1]]
        va = nid(ci + a[0], cj + a[1], ck + a[2])
        vb = nid(ci + b[0], cj + b[1], ck + b[2])
        v111 = nid(ci + 1, cj + 1, ck + 1)
        if sgn > 0:
            cols.append(np.stack([v000, va, vb, v111], axis=1))
        else:
            cols.append(np.stack([v000, vb, va, v111], axis=1))
    # cube-major order: the 6 tets of one cube are consecutive
    tets = np.stack(cols, axis=1).reshape(-1, 4).astype(np.int32)

    facets = _box_facets(nx, ny, nz, nid, X)
    return TetMesh(X=X, tets=tets, facets=facets, shape=(nx, ny, nz), size=(Lx, Ly, Lz))


def _box_facets(nx, ny, nz, nid, X):
    out = []
    dims = (nx, ny, nz)
    for axis in range(3):
        u_ax, v_ax = [a for a in range(3) if a != axis]
        nu, nv = dims[u_ax], dims[v_ax]
        U, V = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
        U = U.ravel(); V = V.ravel()
        for side, outward in ((0, -1.0), (dims[axis], 1.0)):
            def corner(du, dv):
                c = [None, None, None]
                c[axis] = np.full_like(U, side)
                c[u_ax] = U + du
                c[v_ax] = V + dv
                return nid(c[0], c[1], c[2])
            a, b, c, d = corner(0, 0), corner(1, 0), corner(0, 1), corner(1, 1)
            for tri in ((a, b, d), (a, d, c)):
                t = np.stack(tri, axis=1)
                n = np.cross(X[t[:, 1]] - X[t[:, 0]], X[t[:, 2]] - X[t[:, 0]])
                flip = n[:, axis] * outward < 0
                t[flip] = t[flip][:, [0, 2, 1]]
                out.append(t)
    return np.concatenate(out, axis=0).astype(np.int32)


def tet_signed_volumes(X: np.ndarray, tets: np.ndarray) -> np.ndarray:
    a = X[tets[:, 0]]
    return np.einsum("ij,ij->i", X[tets[:, 1]] - a,
                     np.cross(X[tets[:, 2]] - a, X[tets[:, 3]] - a)) / 6.0


def relabel(mesh: TetMesh, perm: np.ndarray) -> TetMesh:
    """Renumber nodes: new node id of old node n is perm[n]."""
    perm = np.asarray(perm, dtype=np.int64)
    X = np.empty_like(mesh.X)
    X[perm] = mesh.X
    return TetMesh(X=X, tets=perm[mesh.tets].astype(np.int32), facets=perm[mesh.facets].astype(np.int32),
                   shape=mesh.shape, size=mesh.size, extra=dict(mesh.extra, perm=perm))


def multicolor_permutation(mesh: TetMesh, colors: int = 8) -> np.ndarray:
    """Node relabelling that groups nodes by colour so no two adjacent nodes share one.

    The 7 edge directions of the Kuhn split are the non-zero 0/1 vectors
    (a, b, c).  With 8 colours, (i + 2j + 4k) mod 8 is proper: the direction
    codes a + 2b + 4c are 1..7.  With 5 colours, (i + j + 2k) mod 5 is proper:
    the codes a + b + 2c are 1..4, never 0 mod 5.  With 4 colours, (i + j + k)
    mod 4 is proper: the codes a + b + c are 1..3 (a tet is a 4-clique, so 4 is
    the fewest possible).  Fewer colours give fewer IC(0) dependency levels.
    Within a colour the natural order is kept.
    Returns perm with new_id = perm[old_id].
    """
    nx, ny, nz = mesh.shape
    n = np.arange(mesh.n_nodes)
    k = n % (nz + 1)
    j = (n // (nz + 1)) % (ny + 1)
    i = n // ((nz + 1) * (ny + 1))
    if colors == 8:
        color = (i + 2 * j + 4 * k) % 8
    elif colors == 5:
        color = (i + j + 2 * k) % 5
    elif colors == 4:
        color = (i + j + k) % 4
    else:
        raise ValueError("multicolour orderings of Kuhn grids: 8, 5 or 4 colours")
    order = np.lexsort((n, color))         # old ids sorted by (colour, old id)
    perm = np.empty_like(order)
    perm[order] = np.arange(mesh.n_nodes)
    return perm


def tiled_color4_permutation(mesh: TetMesh, fixed_nodes: np.ndarray, chunk: int = 64,
                             tile=(4, 8, 8)) -> np.ndarray:
    """The 4-colour numbering ((i + j + k) mod 4, colours in order) with each colour's nodes
    ordered so that every 64-row chunk is a compact tile of the grid, the same tile in every
    colour.

    Block rows are the free nodes in node order, and the library cuts them into chunks of
    `chunk` consecutive rows.  A tile of 4 x 8 x 8 nodes (4 along i, so (i + j + k) mod 4 takes
    each value equally often on every (j, k) line) holds exactly 64 nodes of each colour.
    Within colour c the rows are: a head that completes the chunk left open by the previous
    colour, then one chunk per full tile in tile order, then the nodes of partial tiles sorted
    by tile (chunks of the boundary layers).  Chunk q of the full tiles then covers the same
    tile in every colour, so the grouped SpMV's groups (the four chunks of one tile) hold most
    of their rows' blocks.  Within a colour there are no edges, so the IC(0) factor is the
    one of any other order inside the colours, up to rounding.
    `fixed_nodes` (the nodes that will be constrained, hence not rows) are numbered last.
    Returns perm with new_id = perm[old_id]."""
    nx, ny, nz = mesh.shape
    n = np.arange(mesh.n_nodes)
    k = n % (nz + 1)
    j = (n // (nz + 1)) % (ny + 1)
    i = n // ((nz + 1) * (ny + 1))
    color = (i + j + k) % 4
    free = np.ones(mesh.n_nodes, bool)
    free[np.asarray(fixed_nodes, dtype=np.int64)] = False
    TI, TJ, TK = tile
    i0 = int(i[free].min()) if free.any() else 0
    ti, tj, tk = (i - i0) // TI, j // TJ, k // TK
    # a tile is full when all of its TI*TJ*TK nodes exist and are free
    ntj, ntk = (ny + TJ) // TJ, (nz + TK) // TK
    tile_id = (ti * ntj + tj) * ntk + tk
    cnt = np.bincount(tile_id[free], minlength=int(tile_id.max()) + 1)
    full = free & (cnt[tile_id] == TI * TJ * TK)
    out = []
    rows = 0
    for c in range(4):
        sel = free & (color == c)
        full_nodes = n[sel & full]
        full_nodes = full_nodes[np.lexsort((full_nodes, tile_id[full_nodes]))]
        part = n[sel & ~full]
        part = part[np.lexsort((part, tile_id[part]))]
        head = (-rows) % chunk
        head = min(head, len(part))
        seq = np.concatenate([part[:head], full_nodes, part[head:]])
        out.append(seq)
        rows += len(seq)
    out.append(n[~free])
    order = np.concatenate(out)
    perm = np.empty(mesh.n_nodes, np.int64)
    perm[order] = np.arange(mesh.n_nodes)
    return perm


def gps_relabel(mesh: TetMesh) -> TetMesh:
    """The mesh renumbered by the Gibbs-Poole-Stockmeyer relabeller (od_relabel_gps, host code
    in the native library; MeshRelabeler::getNewLables, core/meshRelabler.cpp:468-498): new node
    labels, and the tets in the relabeller's order (sorted by the sum of their labels)."""
    import ctypes as C
    from . import capi
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32).copy()
    lab = np.zeros(mesh.n_nodes, np.int32)
    rc = capi.lib().od_relabel_gps(mesh.n_nodes, tets.shape[0], tets.shape[1], tets.ctypes.data_as(C.POINTER(C.c_int)),
                                   lab.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != capi.OD_OK:
        raise ValueError(capi.lib().od_last_error().decode())
    perm = lab.astype(np.int64)
    X = np.empty_like(mesh.X)
    X[perm] = mesh.X
    return TetMesh(X=X, tets=tets, facets=perm[mesh.facets].astype(np.int32), shape=mesh.shape, size=mesh.size,
                   extra=dict(mesh.extra, perm=perm))


def plane_fixed_nodes(mesh: TetMesh, axis: int = 0, value: float = 0.0, eps: float = 1e-4) -> np.ndarray:
    """Nodes of boundary facets on the plane X[axis] = value (constrainer.cpp:57-72 rule)."""
    v = np.unique(mesh.facets.ravel())
    sel = np.abs(mesh.X[v, axis] - value) < eps
    return np.sort(v[sel]).astype(np.int32)


def fixed_dofs_from_nodes(nodes: np.ndarray) -> np.ndarray:
    nodes = np.asarray(nodes, dtype=np.int64)
    return np.sort((nodes[:, None] * 3 + np.arange(3)[None, :]).ravel()).astype(np.int32)


def log_uniform(seed: int, stream: int, idx, lo: float, hi: float) -> np.ndarray:
    u = uniform01(seed, stream, idx)
    return np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
