"""Pins of the oracle's integrators against what the mathematics fixes.

- Newmark (full order, linearised per step) reduces to the modal Newmark
  recursion of the reference's LinearNewmark (intergrator/linearNewmark.cpp:
  27-30, 57-79) for a linear material: generalized eigenbasis from scipy, one
  scalar recursion per mode.  The corotational model is linear to first order
  in the load; its second-order part is even in the load, so the odd part
  (run(+f) - run(-f)) / 2 equals the linear response up to third order.
- At beta = 1/4, gamma = 1/2 without load or damping the discrete energy
  1/2 v'Mv + 1/2 d'Kd is conserved exactly by the trapezoidal rule.
- Backward Euler (intergrator/nonlinearImplicitBackwardEuler.h:82-116) on the
  same small load reduces to the dense linear recursion
  (M + h^2 K) v1 = M v0 + h (f - K u0), u1 = u0 + h v1.
A dropped y/z term, a wrong beta/gamma coefficient or a sign error in the
internal force fails one of these.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
from opendeformer_amd import workloads as W

SMALL = 1e-4          # load scale: strains ~1e-4; the odd part is linear to ~1e-8 relative


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def lin():
    w = W.box_small("natural")
    s = O.System(w.X, w.tets, w.fixed_dofs)
    s.set_free_displacements(np.zeros(s.ndof))
    K, _ = s.assemble(w.kind, w.E, w.nu)
    Mv = s.mass(w.rho)
    Kd, Md = s.dense(K), s.dense(Mv)
    lam, phi = sla.eigh(Kd, Md)          # K phi = lam M phi, phi' M phi = I
    f = s.body_force(w.body) * SMALL
    return w, s, Kd, Md, lam, phi, f


def modal_newmark(lam, fq, h, steps, beta=0.25, gamma=0.5):
    q = np.zeros_like(lam); qd = np.zeros_like(lam); qa = fq.copy()   # at rest: q'' = f
    out = []
    for _ in range(steps):
        pred = q + h * qd + (0.5 - beta) * h * h * qa
        qa_new = (fq - lam * pred) / (1.0 + beta * h * h * lam)
        q = pred + beta * h * h * qa_new
        qd = qd + h * ((1.0 - gamma) * qa + gamma * qa_new)
        qa = qa_new
        out.append(q.copy())
    return out


def odd_part(make, steps):
    """Per-step states of run(+) and run(-): their half-difference."""
    a, b = make(+1.0), make(-1.0)
    out = []
    for _ in range(steps):
        a.step(1); b.step(1)
        sa, sb = a.state(), b.state()
        out.append([(x - y) / 2 for x, y in zip(sa, sb)])
    return out


def test_newmark_reduces_to_modal_recursion(oracle, lin):
    w, s, Kd, Md, lam, phi, f = lin
    steps = 30

    def make(sign):
        nm = O.Newmark(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_IC0, p1=1e-3, tol=1e-12)
        nm.set_external(sign * f)
        return nm
    qs = modal_newmark(lam, phi.T @ f, w.dt, steps)
    states = odd_part(make, steps)
    for k in range(steps):
        assert relerr(states[k][0], phi @ qs[k]) < 1e-8, k
    # the full (not antisymmetrised) run differs at first order in the strain: the test resolves it
    nm = make(1.0); nm.step(steps)
    assert relerr(nm.state()[0], phi @ qs[-1]) > 1e-6


def test_newmark_trapezoidal_conserves_discrete_energy(oracle, lin):
    w, s, Kd, Md, lam, phi, f = lin
    rng = np.random.default_rng(5)
    d0 = phi[:, :12] @ rng.standard_normal(12)
    d0 *= 1e-5 / np.abs(d0).max()                          # low modes, 10 um amplitude
    v0 = np.zeros(s.ndof)
    a0 = np.linalg.solve(Md, -Kd @ d0)

    def make(sign):
        nm = O.Newmark(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_IC0, p1=1e-3, tol=1e-13, init_accel=False)
        nm.set_external(np.zeros(s.ndof))
        nm.set_state(sign * d0, sign * v0, sign * a0)
        return nm
    E0 = 0.5 * d0 @ Kd @ d0
    d, v, _ = odd_part(make, 40)[-1]
    E = 0.5 * v @ Md @ v + 0.5 * d @ Kd @ d
    assert abs(E - E0) / E0 < 2e-8
    # and the motion is not trivial: energy moved between kinetic and potential
    assert 0.5 * v @ Md @ v > 1e-3 * E0


def test_backward_euler_reduces_to_dense_linear_recursion(oracle, lin):
    w, s, Kd, Md, lam, phi, f = lin
    h = w.dt
    A = Md + h * h * Kd
    free = np.setdiff1d(np.arange(3 * s.n_nodes), w.fixed_dofs)
    # two systems (the backward-Euler state lives in the system's displacements)
    sys_ = {sg: O.System(w.X, w.tets, w.fixed_dofs) for sg in (1.0, -1.0)}
    be = {}
    for sg, sy in sys_.items():
        be[sg] = O.BackwardEuler(sy, w.kind, w.E, w.nu, w.rho, w.dt, pre=O.PRE_JACOBI, tol=1e-12)
        be[sg].set_external(sg * f)
    u = np.zeros(s.ndof); v = np.zeros(s.ndof)
    for k in range(10):
        for sg in (1.0, -1.0):
            be[sg].step(); be[sg].update_mesh()
        v = np.linalg.solve(A, Md @ v + h * (f - Kd @ u))
        u = u + h * v
        vo = (be[1.0].velocity() - be[-1.0].velocity()) / 2
        uo = (sys_[1.0].node_displacements() - sys_[-1.0].node_displacements()).reshape(-1)[free] / 2
        assert relerr(vo, v) < 1e-8, k
        assert relerr(uo, u) < 1e-8, k
