"""Plane collision damping on the device (od_plane_damping, od_integrator_add_plane_damping)
against the oracle (collision/penaltyDampingForcer.cpp:8-61, collision/intersection.cpp:
12-143, intergrator/nonlinearImplicitBackwardEuler.h:82-101, 122-128), through the C-ABI."""
import numpy as np
import pytest

import oracle as O
from opendeformer_amd import capi
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu

DAMP_RTOL = 1e-13     # damping values: same operations in the same order, contraction off
BE_RTOL = 1e-6        # per-step velocity vs the oracle (PCG to 1e-6)


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(ctx, w):
    import opendeformer_amd as od
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs, w.inst_node_start, w.inst_tet_start)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    return mesh, O.System(w.X, w.tets, w.fixed_dofs)


def contact_pairs(w):
    """A tilted floor under the bottom layer, some pairs repeated, and a second plane on a
    few tets, so entries are summed from several pairs."""
    n = np.array([0.08, -0.05, 1.0]); n /= np.linalg.norm(n)
    planes, tets = W.plane_contacts(w.X, w.tets, n, -0.045)
    planes = np.concatenate([planes, planes[:5], np.tile([[0.0, 0.6, 0.8, -0.09]], (7, 1))])
    tets = np.concatenate([tets, tets[:5], tets[10:17]]).astype(np.int32)
    return planes, tets


@pytest.mark.parametrize("ordering", ["natural", "color5"])
def test_plane_damping_matrix_matches_oracle(ctx, ordering):
    import opendeformer_amd as od
    w = W.box_small(ordering)
    mesh, s = make(ctx, w)
    planes, tets = contact_pairs(w)
    rng = np.random.default_rng(21)
    disp = np.zeros((w.n_nodes, 3)); disp[:, 2] = -0.004 * rng.random(w.n_nodes)
    vel = rng.normal(size=(w.n_nodes, 3)); vel[:, 2] -= 1.0
    disp[w.fixed_nodes] = 0.0; vel[w.fixed_nodes] = 0.0
    s.set_node_displacements(disp.reshape(-1))
    du = od.DeviceArray.from_host(ctx, w.node_to_free(disp))
    dv = od.DeviceArray.from_host(ctx, w.node_to_free(vel))
    print(f"damping: {ordering}, {len(tets)} pairs", flush=True)

    D = od.Matrix(mesh)
    D.add_plane_damping(planes, tets, 300.0, du.ptr, dv.ptr)
    ref = s.plane_damping(planes, tets, 300.0, vel)
    got = D.download()
    assert np.abs(ref).max() > 0
    assert relerr(got, ref) < DAMP_RTOL
    # a second call accumulates, summed in pair order after the first
    D.add_plane_damping(planes, tets, 300.0, du.ptr, dv.ptr)
    ref2 = s.plane_damping(planes, tets, 300.0, vel, values=ref.copy())
    assert relerr(D.download(), ref2) < DAMP_RTOL
    # the same call gives the same bits on a fresh matrix
    D2 = od.Matrix(mesh)
    D2.add_plane_damping(planes, tets, 300.0, du.ptr, dv.ptr)
    assert np.array_equal(D2.download(), got)
    # moving out of the planes adds nothing
    up = od.DeviceArray.from_host(ctx, w.node_to_free(np.tile([0.0, 0.0, 1.0], (w.n_nodes, 1))))
    D3 = od.Matrix(mesh)
    D3.add_plane_damping(planes, tets, 300.0, du.ptr, up.ptr)
    assert not D3.download().any()


def test_plane_damping_rejects_bad_pairs(ctx):
    import ctypes as C
    import opendeformer_amd as od
    w = W.box_small()
    mesh, _ = make(ctx, w)
    D = od.Matrix(mesh)
    planes = np.array([[0.0, 0.0, 1.0, 0.0]])
    bad = np.array([w.n_tets], np.int32)
    rc = capi.lib().od_plane_damping(D.h, 1, capi.dptr(planes), capi.iptr(bad), 1.0, None, None)
    assert rc == capi.OD_ERR_ARG and b"outside the mesh" in capi.lib().od_last_error()
    nm = od.NewmarkIntegrator(mesh, w.dt)
    ok = np.array([0], np.int32)
    rc = capi.lib().od_integrator_add_plane_damping(nm.h, 1, capi.dptr(planes), capi.iptr(ok), 1.0)
    assert rc == capi.OD_ERR_UNSUPPORTED
    assert not D.download().any()
    del C


def test_backward_euler_with_plane_damping_matches_oracle(ctx):
    """The beam sags under gravity onto a tilted floor; each step clears the damping matrix,
    adds the pairs of the tets at or below the floor, steps and moves the mesh."""
    import opendeformer_amd as od
    w = W.box_small()
    mesh, s = make(ctx, w)
    f = s.body_force(w.body)
    be = od.BackwardEulerIntegrator(mesh, w.dt, precond=capi.OD_PRE_JACOBI)
    be.set_external_virtual_work(f)
    ref = O.BackwardEuler(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_JACOBI)
    ref.set_external(f)
    s_free = O.System(w.X, w.tets, w.fixed_dofs)           # the same run without damping
    free_run = O.BackwardEuler(s_free, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_JACOBI)
    free_run.set_external(f)
    n = np.array([0.02, 0.01, 1.0]); n /= np.linalg.norm(n)
    coeff, contacts = 2e6, 0
    for k in range(20):
        x = w.X + s.node_displacements().reshape(-1, 3)
        planes, tets = W.plane_contacts(x, w.tets, n, 0.01)
        contacts += len(tets)
        be.clear_external_damping_force_matrix(); ref.clear_damping()
        be.add_external_damping_force_matrix(planes, tets, coeff); ref.add_plane_damping(planes, tets, coeff)
        assert relerr(be.damping_matrix(), ref.damping()) < 1e-9 or not ref.damping().any(), k
        be.run_one_time_step(); be.update_mesh_vertices_displacements()
        ref.step(); ref.update_mesh()
        free_run.step(); free_run.update_mesh()
        _, v, _ = be.state()
        assert relerr(v, ref.velocity()) < BE_RTOL, k
    assert contacts > 100
    assert relerr(ref.velocity(), free_run.velocity()) > 1e-3
