"""Every IC(0) sweep form gives the same PCG bits, in the mixed form's FP32 sweeps and in the
FP64 sweeps: the lane-per-row kernels with 4, 8 or 16 lanes and the slot kernel with 2, 4, 8
or 16 slots per thread, on meshes whose levels are
consecutive row ranges (colour orderings, where the last sweeps also form the r.z partials),
scattered rows (natural and GPS numbering), several instances in one system, and rows longer
than 16 blocks (a Delaunay mesh).  The solves match a dense solve.
"""
import numpy as np
import pytest

import oracle as O
from opendeformer_amd import capi
from opendeformer_amd import workloads as W
from test_gpu_long_rows import delaunay_workload

pytestmark = pytest.mark.gpu

FORMS = [(8, 0), (4, 0), (16, 0), (8, 2), (8, 4), (8, 8), (8, 16)]


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def solve_every_form(ctx, w, seed, precision=capi.OD_MIXED):
    import opendeformer_amd as od
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs, w.inst_node_start, w.inst_tet_start)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    s = O.System(w.X, w.tets, w.fixed_dofs)
    K, _ = s.assemble(w.kind, w.E, w.nu)
    A = s.mass(w.rho) + 0.25 * w.dt ** 2 * K
    Ad = od.Matrix(mesh); Ad.upload(A); Ad.sync_f32()
    b = np.random.default_rng(seed).standard_normal(mesh.ndof)
    xs = np.linalg.solve(s.dense(A), b) if mesh.ndof <= 6000 else None
    outs = []
    for lanes, slots in FORMS:
        od.mesh_sweep_form(mesh, lanes, slots)
        print(w.name, "sweep form", lanes, slots, "precision", precision, flush=True)
        solver = od.CGSolver(mesh, tol=1e-8, max_iter=3000, precond=capi.OD_PRE_IC0, precond_param=1e-3,
                             precision=precision)
        dx = od.DeviceArray(ctx, mesh.ndof); dx.zero()
        it, _ = solver.solve(Ad, od.DeviceArray.from_host(ctx, b), dx)
        sol = dx.download()
        if xs is not None:
            assert relerr(sol, xs) < 1e-6, (lanes, slots)
        outs.append(((lanes, slots), np.asarray(it).copy(), sol))
    for form, it, sol in outs[1:]:
        assert np.array_equal(it, outs[0][1]) and np.array_equal(sol, outs[0][2]), form
    return mesh


PRECISIONS = [capi.OD_MIXED, capi.OD_FP64]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("ordering", ["natural", "color4", "color4t", "color8", "gps"])
def test_sweep_forms_same_bits_box(ctx, ordering, precision):
    mesh = solve_every_form(ctx, W.box_small(ordering), 31, precision)
    info = mesh.info_raw()
    assert bool(info[12] & 1) == ordering.startswith("color")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_sweep_forms_same_bits_batch(ctx, precision):
    wb = W.batch(n_inst=5, seed=12, tet_lo=300, tet_hi=2500, steps=5)
    solve_every_form(ctx, wb, 32, precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_sweep_forms_same_bits_long_rows(ctx, precision):
    solve_every_form(ctx, delaunay_workload(), 33, precision)
