"""GPU parity of the reference's scalar preconditioners: threshold IC (InCholeskyPreconditioner,
solver/inCholeskyPreconditioner.cpp:8-162) and SAINV LDL^T (InLDLTPreconditioner,
solver/inLDLTPreconditioner.cpp:8-184), factored on the host at every solve and applied by
the device's level-scheduled sweeps, against the oracle at non-zero tolerances.

Bars: the apply z = P^-1 r within 1e-13 relative (the host factor equals the oracle's bit for
bit, tests/test_scalar_precond.py; the device sweeps keep the reference's operation order);
PCG iteration counts within 1 of the oracle's; Newmark displacements within 1e-6 per step.
"""
import numpy as np
import pytest

import oracle as O
from opendeformer_amd import capi
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [(capi.OD_PRE_ICT, O.PRE_ICT, 1e-3, 1.0), (capi.OD_PRE_ICT, O.PRE_ICT, 2e-2, 0.5),
         (capi.OD_PRE_LDLT, O.PRE_LDLT, 1e-3, 1e-2), (capi.OD_PRE_LDLT, O.PRE_LDLT, 1e-2, 5e-2)]


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def setup(ctx, ordering):
    import opendeformer_amd as od
    w = W.box_small(ordering)
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    s = O.System(w.X, w.tets, w.fixed_dofs)
    M = s.mass(w.rho)
    s.set_free_displacements(np.random.default_rng(4).standard_normal(s.ndof) * 0.01)
    K, _ = s.assemble(w.kind, w.E, w.nu)
    return w, mesh, s, M + 0.25 * w.dt ** 2 * K


@pytest.mark.parametrize("ordering", ["natural", "color4"])
@pytest.mark.parametrize("pre,opre,p1,p2", CASES)
def test_scalar_precond_apply_and_pcg_match_oracle(ctx, ordering, pre, opre, p1, p2):
    import opendeformer_amd as od
    w, mesh, s, A = setup(ctx, ordering)
    Ad = od.Matrix(mesh); Ad.upload(A)
    solver = od.CGSolver(mesh, tol=1e-6, max_iter=5000, precond=pre, precond_param=p1)
    solver.set_precond_params(p1, p2)
    r = np.random.default_rng(3).standard_normal(mesh.ndof)
    z = od.DeviceArray(ctx, mesh.ndof)
    solver.precond_apply(Ad, od.DeviceArray.from_host(ctx, r), z)
    zo = s.precond_apply(A, r, opre, p1, p2)
    assert relerr(z.download(), zo) < 1e-13
    rng = np.random.default_rng(7)
    b = rng.standard_normal(mesh.ndof); x0 = rng.standard_normal(mesh.ndof) * 1e-3
    dx = od.DeviceArray.from_host(ctx, x0)
    it, rr = solver.solve(Ad, od.DeviceArray.from_host(ctx, b), dx)
    xo, ito, rro = s.cg(A, b, x0, tol=1e-6, max_iter=5000, pre=opre, p1=p1, p2=p2)
    print(f"{ordering} pre {pre} ({p1}, {p2}): iterations GPU {int(it[0])} oracle {ito}", flush=True)
    assert abs(int(it[0]) - ito) <= 1
    assert rr[0] <= 1e-6 and rro <= 1e-6
    assert relerr(dx.download(), xo) < 1e-5


@pytest.mark.parametrize("pre,opre,p1,p2", [CASES[0], CASES[2]])
def test_newmark_with_scalar_precond_matches_oracle(ctx, pre, opre, p1, p2):
    import opendeformer_amd as od
    w = W.box_small("color4")
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    s = O.System(w.X, w.tets, w.fixed_dofs)
    f = s.body_force(w.body)
    nm = od.NewmarkIntegrator(mesh, w.dt, precond=pre, precond_param=p1)
    nm.set_precond_params(p1, p2)
    nm.set_external_virtual_work(f)
    ref = O.Newmark(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=opre, p1=p1, p2=p2)
    ref.set_external(f)
    for k in range(20):
        nm.run_one_time_step()
        oi, _, _ = ref.step(1)
        assert relerr(nm.state()[0], ref.state()[0]) < 1e-6, k
        assert abs(int(nm.stats()[0][0]) - int(oi[0])) <= 1, k


def test_scalar_precond_is_fp64_only(ctx):
    import opendeformer_amd as od
    w = W.box_small()
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    for pre in (capi.OD_PRE_ICT, capi.OD_PRE_LDLT):
        with pytest.raises(capi.ODError, match="FP64 only"):
            od.CGSolver(mesh, precond=pre, precond_param=1e-3, precision=capi.OD_MIXED)
