"""Whole-step graphs (graph mode of the Newmark integrator): one captured graph per step
(element pass, gather, factor, a fixed plan of PCG iterations or refinement sweeps with
device-side exits, corrector for the instances the graph finished).  Every step must give the
eager path's bits, per instance, including steps whose plan falls short (the instances left are
stepped again eagerly), masked instances and the mixed form; the graph must actually run.
"""
import numpy as np
import pytest

from opendeformer_amd import capi
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu


def build(ctx, w, precision, graph, pre=capi.OD_PRE_IC0):
    import opendeformer_amd as od
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs, w.inst_node_start, w.inst_tet_start)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    f = mesh.body_force(w.body)
    nm = od.NewmarkIntegrator(mesh, w.dt, precond=pre, precision=precision)
    nm.set_graph(graph, whole_step=True)
    nm.set_external_virtual_work(f)
    return mesh, nm, f


def run(ctx, w, precision, graph, steps, hook=None, pre=capi.OD_PRE_IC0):
    mesh, nm, f = build(ctx, w, precision, graph, pre)
    out = []
    for k in range(steps):
        if hook:
            hook(k, mesh, nm, f)
        nm.run_one_time_step()
        d, v, a = nm.state()
        it, rr, outer, _ = nm.stats()
        out.append((d.copy(), v.copy(), a.copy(), it.copy(), rr.copy(), outer))
    return out, nm.graph_info()


def same(e, g):
    for k, (x, y) in enumerate(zip(e, g)):
        for a, b in zip(x[:5], y[:5]):
            assert np.array_equal(a, b), f"step {k}"
        assert x[5] == y[5], f"step {k}: refinement sweeps"


@pytest.mark.parametrize("precision", [capi.OD_FP64, capi.OD_MIXED])
def test_step_graph_single_mesh_equals_eager(ctx, precision):
    w = W.box_small("color4")
    e, _ = run(ctx, w, precision, False, 24)
    g, info = run(ctx, w, precision, True, 24)
    print(f"precision {precision}: {info}", flush=True)
    same(e, g)
    assert info["graph_steps"] == 23 and info["captures"] <= 4


@pytest.mark.parametrize("precision", [capi.OD_FP64, capi.OD_MIXED])
def test_step_graph_batch_equals_eager(ctx, precision):
    wb = W.batch(n_inst=8, seed=21, tet_lo=300, tet_hi=6000, steps=20)
    e, _ = run(ctx, wb, precision, False, 20)
    g, info = run(ctx, wb, precision, True, 20)
    print(f"precision {precision}: {info}", flush=True)
    same(e, g)
    assert info["graph_steps"] == 19


@pytest.mark.parametrize("precision", [capi.OD_FP64, capi.OD_MIXED])
def test_step_graph_short_plan_completes_eagerly(ctx, precision):
    """Material 100 times stiffer from step 6 on: that step needs more iterations than the
    plan drawn from step 5, the graph leaves instances, the eager remainder completes them,
    bits unchanged; the plans of the later steps follow the new counts."""
    wb = W.batch(n_inst=5, seed=22, tet_lo=300, tet_hi=4000, steps=12)

    def hook(k, mesh, nm, f):
        if k == 6:
            mesh.set_material(wb.kind, wb.E * 100.0, wb.nu, wb.rho)

    e, _ = run(ctx, wb, precision, False, 12, hook)
    g, info = run(ctx, wb, precision, True, 12, hook)
    print(f"precision {precision}: {info}; iterations {[int(x[3].max()) for x in g]}", flush=True)
    same(e, g)
    assert info["fallbacks"] >= 1 and info["graph_steps"] >= 9


def test_step_graph_with_inactive_instances(ctx):
    wb = W.batch(n_inst=6, seed=23, tet_lo=300, tet_hi=3000, steps=10)
    mask = np.array([1, 0, 1, 1, 0, 1], np.uint8)

    def hook(k, mesh, nm, f):
        if k == 3:
            nm.set_active(mask)
        if k == 7:
            nm.set_active(None)

    e, _ = run(ctx, wb, capi.OD_FP64, False, 10, hook)
    g, info = run(ctx, wb, capi.OD_FP64, True, 10, hook)
    same(e, g)
    assert info["graph_steps"] == 9


def test_step_graph_jacobi(ctx):
    w = W.box_small("natural")
    e, _ = run(ctx, w, capi.OD_FP64, False, 10, pre=capi.OD_PRE_JACOBI)
    g, info = run(ctx, w, capi.OD_FP64, True, 10, pre=capi.OD_PRE_JACOBI)
    same(e, g)
    assert info["graph_steps"] == 9


@pytest.mark.parametrize("precision", [capi.OD_FP64, capi.OD_MIXED])
def test_step_graphs_captured_per_plan_and_sweep_form(ctx, precision):
    """Graph mode: the first (eager) step captures the PCG iteration block once for the
    integrator's buffers; later steps replay whole-step graphs, captured once per iteration plan
    and replayed while the plan holds; a change of the sweep form (baked into the captured
    launches) makes a new step graph."""
    import opendeformer_amd as od
    w = W.box_small("color4")
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    nm = od.NewmarkIntegrator(mesh, w.dt, precond=capi.OD_PRE_IC0, precision=precision)
    nm.set_graph(True, whole_step=True)
    nm.set_external_virtual_work(mesh.body_force(w.body))
    nm.run_time_steps(8)
    assert nm.solver_info()["graph_captures"] == 1
    g0 = nm.graph_info()
    assert g0["graph_steps"] == 7 and 1 <= g0["captures"] <= 3
    od.mesh_sweep_form(mesh, 8, 4 if precision == capi.OD_MIXED else 0)
    nm.run_time_steps(3)
    g1 = nm.graph_info()
    assert g1["graph_steps"] == 10 and 1 <= g1["captures"] - g0["captures"] <= 2
    assert nm.solver_info()["graph_captures"] == 1
