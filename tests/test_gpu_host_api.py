"""The C++ host classes (include/od_gpu.hpp), used the way host code written against the
reference's interfaces uses them (TetMesh, Constrainer, NodeIndexer, a material's
getMatrixStructure / getMatrixIndicesPerElement / generateMatrixAndVirtualWorks, SpMDV, a
CGSolver with an IC(0) Preconditioner, a Forcer, a Simulator stepping a GpuNewmark, a
GpuBackwardEuler with a plane collision damping forcer), against the oracle.

tests/cpp/host_api_check is built by the package build; this test writes its
input file, runs it and compares every output it writes.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from opendeformer_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_api_check")


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_host_classes_match_oracle(tmp_path):
    assert os.path.exists(BIN), "host_api_check is not built (run the package build)"
    w = W.box_small("natural")
    s = O.System(w.X, w.tets, w.fixed_dofs)
    n, nn = s.ndof, w.X.shape[0]
    free = np.setdiff1d(np.arange(3 * nn), w.fixed_dofs)
    rng = np.random.default_rng(21)
    u_free = rng.standard_normal(n) * 0.01
    u_node = np.zeros(3 * nn); u_node[free] = u_free
    x = rng.standard_normal(n); y0 = rng.standard_normal(n); b = rng.standard_normal(n)
    g = s.body_force(w.body)
    steps = 12
    rho, E, nu = float(w.rho[0]), float(w.E[0]), float(w.nu[0])
    # candidate pairs: the tets within 5 cm of a tilted floor 1 cm under the beam (the
    # sagging beam reaches it after a few steps; pairs not yet touching add nothing)
    nrm = np.array([0.02, 0.01, 1.0]); nrm /= np.linalg.norm(nrm)
    planes, ptets = W.plane_contacts(w.X, w.tets, nrm, 0.01 - 0.05)
    planes[:, 3] = 0.01
    coeff = 2e6
    inp = tmp_path / "in.bin"; out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([nn, w.tets.shape[0], len(w.fixed_dofs), n, steps], np.int32).tofile(f)
        w.X.astype(np.float64).ravel().tofile(f)
        w.tets.astype(np.int32).ravel().tofile(f)
        np.asarray(w.fixed_dofs, np.int32).tofile(f)
        np.array([rho, E, nu, w.dt], np.float64).tofile(f)
        for v in (u_node, x, y0, b):
            np.asarray(v, np.float64).tofile(f)
        np.array([len(ptets)], np.int32).tofile(f)
        planes.astype(np.float64).ravel().tofile(f)
        ptets.astype(np.int32).tofile(f)
        np.array([coeff], np.float64).tofile(f)
    print("host_api_check on box_small", flush=True)
    r = subprocess.run([BIN, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(out, "rb").read()
    nv = int(np.frombuffer(raw[:8], np.int64)[0])
    off = 8

    def take(count, dt=np.float64):
        nonlocal off
        a = np.frombuffer(raw, dt, count, off); off += a.nbytes
        return a

    K = take(nv); vws = take(n); y = take(n); sol = take(n); iters = int(take(1, np.int32)[0])
    z_ic0 = take(n); z_ict = take(n)
    g_host = take(n); d_node = take(3 * nn); v_node = take(3 * nn); pos = take(3 * nn, np.float32)
    v_be = take(3 * nn)
    assert relerr(g_host, g) < 1e-13        # Forcer::addBodyForce + getVirtualWorks

    s.set_free_displacements(u_free)
    Kc, fc = s.assemble(w.kind, w.E, w.nu)
    assert relerr(K, Kc) < 1e-11
    assert relerr(vws, fc) < 1e-10
    assert relerr(y, s.spmdv(Kc, x, y0.copy())) < 1e-12
    A = s.mass(w.rho) + 0.25 * w.dt ** 2 * Kc
    xs = np.linalg.solve(s.dense(A), b)
    assert 0 < iters < 5000
    assert relerr(sol, xs) < 1e-8
    assert relerr(z_ic0, s.precond_apply(A, b, O.PRE_IC0, 1e-3)) < 1e-10
    assert relerr(z_ict, s.precond_apply(A, b, O.PRE_ICT, 1e-3, 1.0)) < 1e-13

    ref = O.Newmark(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_IC0, p1=1e-3)
    ref.set_external(g)
    ref.step(steps)
    do, vo, _ = ref.state()
    assert relerr(d_node[free], do) < 1e-6
    assert relerr(v_node[free], vo) < 1e-6
    assert np.all(d_node[np.asarray(w.fixed_dofs)] == 0)
    # Simulator::getVertexPositions: float X + u
    d_full = np.zeros(3 * nn); d_full[free] = do
    np.testing.assert_allclose(pos, (w.X.ravel() + d_full).astype(np.float32), rtol=1e-6, atol=1e-6)

    s2 = O.System(w.X, w.tets, w.fixed_dofs)
    be = O.BackwardEuler(s2, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_IC0, p1=1e-3)
    be.set_external(g)
    touched = 0
    for _ in range(steps):
        be.clear_damping()
        be.add_plane_damping(planes, ptets, coeff)
        touched += int(be.damping().any())
        be.step(); be.update_mesh()
    assert touched >= 3
    assert relerr(v_be[free], be.velocity()) < 1e-6
    assert np.all(v_be[np.asarray(w.fixed_dofs)] == 0)
