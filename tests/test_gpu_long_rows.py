"""Block rows longer than 16 blocks (the SpMV's row-loop path) on an unstructured mesh: a
Delaunay tetrahedralisation of random points, whose nodes have up to about 30 neighbours.

Every product shape (chunks per workgroup, lanes per row, and the FP32 slot-value form) gives
the same PCG bits, the FP64 products match the oracle's, and the solve matches a dense one.
"""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import oracle as O
from opendeformer_amd import capi
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def delaunay_workload(n=600, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 3)) * np.array([1.0, 0.5, 0.5])
    tets = Delaunay(X).simplices.astype(np.int32)
    a, b, c, d = (X[tets[:, k]] for k in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    tets = tets[np.abs(vol) > 1e-6].copy()            # drop slivers of the hull
    vol = vol[np.abs(vol) > 1e-6]
    flip = vol < 0
    tets[flip, 2], tets[flip, 3] = tets[flip, 3].copy(), tets[flip, 2].copy()
    used = np.unique(tets)
    remap = -np.ones(n, np.int32); remap[used] = np.arange(len(used), dtype=np.int32)
    X, tets = X[used], remap[tets]
    fixed = np.nonzero(X[:, 0] < 0.08)[0].astype(np.int32)
    m = len(tets)
    return W.Workload("delaunay", X, tets, np.zeros((0, 3), np.int32), fixed, np.zeros(m, np.int32),
                      np.full(m, 1e5), np.full(m, 0.4), np.full(m, 1000.0), 1.0 / 60, 1)


def longest_row(w):
    """Blocks in the longest block row: a free node and its free neighbours."""
    free = np.ones(len(w.X), bool); free[w.fixed_nodes] = False
    pairs = {(int(a), int(b)) for t in w.tets for a in t for b in t if free[a] and free[b]}
    deg = np.bincount([a for a, _ in pairs], minlength=len(w.X))
    return int(deg.max())


def test_long_rows_every_shape_same_bits_and_oracle(ctx):
    import opendeformer_amd as od
    w = delaunay_workload()
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    s = O.System(w.X, w.tets, w.fixed_dofs)
    assert longest_row(w) > 16, "the mesh must exercise the row-loop path"
    Mv = s.mass(w.rho)
    K, _ = s.assemble(w.kind, w.E, w.nu)
    A = Mv + 0.25 * w.dt ** 2 * K
    Ad = od.Matrix(mesh); Ad.upload(A)
    rng = np.random.default_rng(23)
    x = rng.standard_normal(mesh.ndof)
    dy = od.DeviceArray(ctx, mesh.ndof)
    od.spmdv(Ad, od.DeviceArray.from_host(ctx, x), dy, accumulate=False)
    assert relerr(dy.download(), s.spmdv(A, x, np.zeros(mesh.ndof))) < 1e-12
    b = rng.standard_normal(mesh.ndof)
    xs = np.linalg.solve(s.dense(A), b)
    shapes = [((r, ls), (r, ls)) for r, ls in ((1, 16), (2, 8), (4, 4))] + [((1, 16), (k, 0)) for k in (1, 2, 4)]
    outs = []
    for v64, v32 in shapes:
        od.mesh_spmv_variant(mesh, v64, v32)
        for prec in (capi.OD_FP64, capi.OD_MIXED):
            if prec == capi.OD_MIXED:
                Ad.sync_f32()
            print("long rows: shape", v64 if prec == capi.OD_FP64 else v32, "precision", prec, flush=True)
            solver = od.CGSolver(mesh, tol=1e-8, max_iter=3000, precond=capi.OD_PRE_IC0, precond_param=1e-3,
                                 precision=prec)
            dx = od.DeviceArray(ctx, mesh.ndof); dx.zero()
            it, _ = solver.solve(Ad, od.DeviceArray.from_host(ctx, b), dx)
            sol = dx.download()
            assert relerr(sol, xs) < 1e-6
            outs.append((v64 if prec == capi.OD_FP64 else v32, prec, int(it[0]), sol))
    for prec in (capi.OD_FP64, capi.OD_MIXED):
        sel = [o for o in outs if o[1] == prec]
        for o in sel[1:]:
            assert o[2] == sel[0][2] and np.array_equal(o[3], sel[0][3]), (prec, o[0])
