"""Pins of the oracle's corotational plastic material (material/corotationalPlastic.cpp).

- Without yielding, its stiffness equals the corotational one exactly and its forces
  (from the stress R (lambda tr(e) I + 2 mu e), e = S - I) equal the corotational forces
  R K0 (R^T x - X) to rounding: for a linear tet both are V B^T D B applied in the rotated
  frame.
- With creep ratio 1 and no cap, one evaluation moves the plastic strain so that the elastic
  strain's norm (the reference's norm, 3 (e00^2 + e01^2 + e02^2), :159) becomes the yield
  criterion exactly; a second evaluation at the same deformation leaves it unchanged.
- The cap holds the plastic strain's norm (same reference norm, :172) at the maximum.
"""
import numpy as np
import pytest

import oracle as O
from opendeformer_amd import workloads as W


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def quirk_norm(e):
    # sum over i, j < 3 of e[i]^2 with e the flattened 3x3 strain: 3 * (e[0]^2 + e[1]^2 + e[2]^2)
    f = np.asarray(e).reshape(-1, 9)
    return np.sqrt(3.0 * (f[:, 0] ** 2 + f[:, 1] ** 2 + f[:, 2] ** 2))


def polar_S(F):
    U, s, Vt = np.linalg.svd(F)
    R = U @ Vt
    return R.T @ F


@pytest.fixture()
def deformed():
    w = W.box_small()
    s = O.System(w.X, w.tets, w.fixed_dofs)
    rng = np.random.default_rng(31)
    s.set_free_displacements(rng.standard_normal(s.ndof) * 0.02)
    return w, s


def test_plastic_without_yield_is_corotational(oracle, deformed):
    w, s = deformed
    Kc, fc = s.assemble(0, w.E, w.nu)
    s.set_plastic(np.inf, 0.0, np.inf)
    Kp, fp = s.assemble(2, w.E, w.nu)
    assert np.array_equal(Kp, Kc)
    assert relerr(fp, fc) < 1e-10
    assert not s.plastic_strain().any()


def elastic_strains(w, s):
    """sym(S) - I - plastic strain per tet, S from an SVD polar of F (numpy)."""
    nodes = s.node_displacements() + w.X
    out = []
    ep = s.plastic_strain()
    for e, t in enumerate(w.tets):
        Dm = (w.X[t[1:]] - w.X[t[0]]).T
        Ds = (nodes[t[1:]] - nodes[t[0]]).T
        F = Ds @ np.linalg.inv(Dm)
        S = polar_S(F)
        out.append(0.5 * (S + S.T) - np.eye(3) - ep[e])
    return np.array(out)


def test_full_creep_returns_to_the_yield_surface(oracle, deformed):
    w, s = deformed
    e0 = elastic_strains(w, s)
    n0 = quirk_norm(e0)
    yld = float(np.median(n0))                 # about half of the tets yield
    s.set_plastic(yld, 1.0, np.inf)
    s.assemble(2, w.E, w.nu)
    ep1 = s.plastic_strain().copy()
    yielded = n0 > yld
    assert yielded.sum() > 100 and (~yielded).sum() > 100
    assert not ep1[~yielded].any()
    n1 = quirk_norm(elastic_strains(w, s))
    np.testing.assert_allclose(n1[yielded], yld, rtol=1e-9)
    # a second evaluation at the same deformation: on the yield surface, no further creep
    s.assemble(2, w.E, w.nu)
    np.testing.assert_allclose(s.plastic_strain(), ep1, rtol=0, atol=1e-12 * np.abs(ep1).max())


def test_cap_holds_plastic_strain_at_the_maximum(oracle, deformed):
    w, s = deformed
    n0 = quirk_norm(elastic_strains(w, s))
    cap = 1e-4
    s.set_plastic(0.0, 1.0, cap)
    s.assemble(2, w.E, w.nu)
    pn = quirk_norm(s.plastic_strain())
    big = pn > 0
    assert big.sum() > 100
    np.testing.assert_allclose(pn[n0 > cap * 1.01], cap, rtol=1e-9)
    assert np.all(pn <= cap * (1 + 1e-12))
