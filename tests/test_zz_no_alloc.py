"""od_pcg_solve_noalloc (every temporary in the library's fixed device scratch) against
od_pcg_solve with the same preconditioner: bit-identical iterates, iteration counts and
residuals, on two calls in a row, the second on a smaller system (the scratch keeps the first
call's contents, which the second must not see)."""
import numpy as np
import pytest

import oracle as O
from opendeformer_amd import capi
from opendeformer_amd import mesh as M
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu


def system(ctx, w):
    import opendeformer_amd as od
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    s = O.System(w.X, w.tets, w.fixed_dofs)
    Mv = s.mass(w.rho)
    s.set_free_displacements(np.random.default_rng(4).standard_normal(s.ndof) * 0.01)
    K, _ = s.assemble(w.kind, w.E, w.nu)
    A = od.Matrix(mesh)
    A.upload(Mv + 0.25 * w.dt ** 2 * K)
    return mesh, A


@pytest.mark.parametrize("grouped", [False, True], ids=["tuned", "grouped"])
@pytest.mark.parametrize("pre", [capi.OD_PRE_JACOBI, capi.OD_PRE_NONE])
def test_noalloc_solve_equals_pcg_solve_on_two_calls(ctx, pre, grouped):
    """With grouped, the mesh runs the grouped SpMV form, whose values the allocation-free
    solve cannot hold: it runs the chunk-descriptor shape instead (the same bits)."""
    import opendeformer_amd as od
    small = W._single("box_8x3x3", M.box_mesh(8, 3, 3, 0.4, 0.15, 0.15, seed=1), 0, 1e5, 0.45, 1.0 / 60, 1, "color4")
    for call, w in enumerate([W.box_small("color4t" if grouped else "color4"), small]):
        mesh, A = system(ctx, w)
        if grouped:
            od.mesh_spmv_variant(mesh, (8, -1), (8, -1))
        rng = np.random.default_rng(10 + call)
        b = od.DeviceArray.from_host(ctx, rng.standard_normal(mesh.ndof))
        x0 = rng.standard_normal(mesh.ndof) * 1e-3
        print(f"call {call}: {mesh.ndof} DOF, preconditioner {pre}", flush=True)
        solver = od.CGSolver(mesh, tol=1e-8, max_iter=3000, precond=pre)
        xa = od.DeviceArray.from_host(ctx, x0)
        ita, rra = solver.solve(A, b, xa)
        xb = od.DeviceArray.from_host(ctx, x0)
        itb, rrb = od.pcg_solve_noalloc(A, b, xb, precond=pre, tol=1e-8, max_iter=3000)
        assert ita[0] > 3 and rra[0] <= 1e-8
        assert np.array_equal(ita, itb) and np.array_equal(rra, rrb)
        assert np.array_equal(xa.download(), xb.download())


def test_noalloc_solve_rejects_what_it_cannot_hold(ctx):
    import opendeformer_amd as od
    mesh, A = system(ctx, W.box_small())
    b = od.DeviceArray(ctx, mesh.ndof); b.zero()
    x = od.DeviceArray(ctx, mesh.ndof); x.zero()
    with pytest.raises(capi.ODError, match="OD_PRE_NONE or OD_PRE_JACOBI"):
        od.pcg_solve_noalloc(A, b, x, precond=capi.OD_PRE_IC0)
