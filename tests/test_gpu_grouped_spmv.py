"""The grouped symmetric SpMV (lanes -1) on the GPU, through the C-ABI.

The grouped form reads each upper block stored for a group once and feeds the column row of a
block inside the group through LDS; every slot term and the 16-lane slot tree are those of the
chunk-descriptor forms.  So on every mesh where it runs, FP64 and mixed PCG solves must give the
descriptor form's bits and iteration counts (one instance, a batch with per-instance
convergence, the tiled and the plain 4-colour numberings, the natural numbering), and Newmark
on the tiled 4-colour numbering must follow the FP64 oracle within 1e-6 per step, as every
other numbering does.
"""
import numpy as np
import pytest

import oracle as O
from opendeformer_amd import capi
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(ctx, w):
    import opendeformer_amd as od
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs, w.inst_node_start, w.inst_tet_start)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    return mesh


def newmark_system(w, seed=4):
    """M + dt^2/4 K at a random displacement, from the oracle, and the oracle system."""
    s = O.System(w.X, w.tets, w.fixed_dofs)
    Mv = s.mass(w.rho)
    s.set_free_displacements(np.random.default_rng(seed).standard_normal(s.ndof) * 0.01)
    K, _ = s.assemble(w.kind, w.E, w.nu)
    return Mv + 0.25 * w.dt ** 2 * K, s


CASES = [("small_color4t", lambda: W.box_small("color4t")),
         ("small_natural", lambda: W.box_small("natural")),
         ("mid_color4t", lambda: W.box_mid_neo("color4t")),
         ("mid_color4", lambda: W.box_mid_neo("color4")),
         ("batch12_color4t", lambda: W.batch(12, seed=4, tet_lo=500, tet_hi=20000, ordering="color4t"))]


@pytest.mark.parametrize("name,make_w", CASES, ids=[c[0] for c in CASES])
def test_grouped_pcg_bits_equal_descriptor_form(ctx, name, make_w):
    import opendeformer_amd as od
    w = make_w()
    print(f"grouped PCG bits: {name}", flush=True)
    mesh = make(ctx, w)
    table, _, _ = od.mesh_autotune(mesh)
    assert all((p, -1, k) in table for p in ("fp64", "fp32") for k in (4, 8))
    Av, _ = newmark_system(w)
    A = od.Matrix(mesh); A.upload(Av)
    b = np.random.default_rng(11).standard_normal(mesh.ndof)
    outs = {}
    for form in ("descriptor", "grouped4", "grouped8"):
        v = ((1, 8), (1, 8)) if form == "descriptor" else ((int(form[-1]), -1), (int(form[-1]), -1))
        od.mesh_spmv_variant(mesh, *v)
        for prec in (capi.OD_FP64, capi.OD_MIXED):
            if prec == capi.OD_MIXED:
                A.sync_f32()
            solver = od.CGSolver(mesh, tol=1e-8, max_iter=3000, precond=capi.OD_PRE_IC0, precond_param=1e-3,
                                 precision=prec)
            dx = od.DeviceArray(ctx, mesh.ndof); dx.zero()
            it, _ = solver.solve(A, od.DeviceArray.from_host(ctx, b), dx)
            outs[(form, prec)] = (np.asarray(it).copy(), dx.download())
    for prec in (capi.OD_FP64, capi.OD_MIXED):
        ia, xa = outs[("descriptor", prec)]
        for form in ("grouped4", "grouped8"):
            ib, xb = outs[(form, prec)]
            assert np.array_equal(ia, ib), (prec, form)
            assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64)), (prec, form)


def test_newmark_tiled_colour4_grouped_matches_oracle(ctx):
    """100 Newmark steps on the tiled 4-colour numbering with the grouped product in both
    precisions' solves: FP64 within 1e-6 per step of the FP64 oracle, PCG iterations within 1."""
    import opendeformer_amd as od
    w = W.box_small("color4t")
    mesh = make(ctx, w)
    od.mesh_spmv_variant(mesh, (8, -1), (4, -1))
    s = O.System(w.X, w.tets, w.fixed_dofs)
    f = s.body_force(w.body)
    nm = od.NewmarkIntegrator(mesh, w.dt, precond=capi.OD_PRE_IC0)
    nm.set_graph(True)
    nm.set_external_virtual_work(f)
    ref = O.Newmark(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_IC0, p1=1e-3)
    ref.set_external(f)
    for step in range(w.steps):
        nm.run_one_time_step()
        it_o, _, _ = ref.step(1)
        d = nm.state()[0]
        do = ref.state()[0]
        err = np.linalg.norm(d - do) / np.linalg.norm(do)
        assert err < 1e-6, (step, err)
        assert abs(int(nm.stats()[0][0]) - int(it_o[0])) <= 1, step


# launch forms: (SpMV shapes FP64, FP32), sweep form (lanes, slots) or None (the mesh's own tuning)
FORMS = [("descriptor", ((1, 16), (1, 16)), (16, 0)),
         ("grouped8_4", ((8, -1), (4, -1)), (8, 0)),
         ("grouped4_8", ((4, -1), (8, -1)), (4, 4)),
         ("slotvalues", ((2, 8), (4, 0)), (8, 2)),
         ("tuned", None, None)]


@pytest.mark.parametrize("precision", [capi.OD_FP64, capi.OD_MIXED], ids=["fp64", "mixed"])
def test_newmark_forms_same_bits_on_served_pool_objects(ctx, precision):
    """Newmark on the served stream's pool objects (tiled 4-colour numbering, 1k-50k tets) with
    an impulse: every SpMV and sweep form, set by the caller or tuned on the mesh, gives the
    bits of the plain descriptor form, step by step."""
    import opendeformer_amd as od
    pool = W.served_pool(8, seed=5, tet_lo=1000, tet_hi=50000)
    rng = np.random.default_rng(3)
    for i, w in enumerate(pool):
        f = np.zeros((w.n_nodes, 3))
        f[rng.integers(0, w.n_nodes, 8)] = rng.standard_normal((8, 3)) * 20
        imp = w.node_to_free(f)
        ref = None
        for name, spmv, sweep in FORMS:
            print(f"pool object {i} ({w.n_tets} tets), form {name}", flush=True)
            mesh = make(ctx, w)
            if spmv is not None:
                od.mesh_spmv_variant(mesh, *spmv)
                od.mesh_sweep_form(mesh, *sweep)
            nm = od.NewmarkIntegrator(mesh, w.dt, precond=capi.OD_PRE_IC0, precision=precision, tol=1e-6)
            nm.set_external_virtual_work(mesh.body_force(w.body))
            nm.set_impulse(imp)
            states = []
            for _ in range(6):
                nm.run_one_time_step()
                states.append((nm.state()[0].copy(), int(nm.stats()[0][0])))
            del nm, mesh
            if ref is None:
                ref = states
                continue
            for step, ((da, ia), (db, ib)) in enumerate(zip(ref, states)):
                assert ia == ib and np.array_equal(da.view(np.uint64), db.view(np.uint64)), (i, name, step)
