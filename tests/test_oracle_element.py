"""Pins the oracle's element kernels with closed forms and finite differences.

Corotational: element/tetelement.cpp:501-651; neo-Hookean: element/tetelement.cpp:253-499,
material/invertibleHyperelastic.cpp:116-265, material/invertibleNeoHookean.cpp:11-20.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from opendeformer_amd import workloads as W

E, NU = 1e5, 0.45
LAM = NU * E / ((1 + NU) * (1 - 2 * NU))
MU = E / (2 * (1 + NU))


def tet(rng, scale=0.1):
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * scale
    return X + rng.uniform(-0.15, 0.15, (4, 3)) * scale


def unpack78(sub):
    K = np.zeros((12, 12))
    k = 0
    for c in range(12):
        for r in range(c, 12):
            K[r, c] = K[c, r] = sub[k]
            k += 1
    return K


def corot(O, X, x, thr=0.0):
    sub = np.zeros(78); f = np.zeros(12); R = np.zeros(9)
    O.lib().oro_corot_element(O._d(O.f64(X.ravel())), O._d(O.f64(x.ravel())), E, NU, thr, O._d(sub), O._d(f), O._d(R))
    return unpack78(sub), f, R.reshape(3, 3)


def neo(O, X, x, thr=0.1):
    sub = np.zeros(78); f = np.zeros(12)
    O.lib().oro_neo_element(O._d(O.f64(X.ravel())), O._d(O.f64(x.ravel())), E, NU, thr, O._d(sub), O._d(f))
    return unpack78(sub), f


def k0(O, X):
    K0 = np.zeros(90)
    O.lib().oro_corot_k0(O._d(O.f64(X.ravel())), E, NU, O._d(K0))
    K = np.zeros((12, 12))
    i = 0
    for a in range(4):
        for b in range(a, 4):
            blk = K0[i:i + 9].reshape(3, 3)
            K[3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
            K[3 * b:3 * b + 3, 3 * a:3 * a + 3] = blk.T
            i += 9
    return K


def grad_n(X):
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    G = np.linalg.inv(Dm)                  # rows: gradients of N1..N3
    g = np.vstack([-G.sum(axis=0), G])
    return g, abs(np.linalg.det(Dm)) / 6.0


def test_k0_rigid_modes_and_patch_test(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20):
        X = tet(rng)
        K = k0(oracle, X)
        np.testing.assert_allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
        ev = np.linalg.eigvalsh(K)
        assert (np.abs(ev[:6]) < 1e-9 * ev[-1]).all() and ev[6] > 1e-6 * ev[-1]
        for a in range(4):                 # sum_b K0_ab = 0 (translations)
            np.testing.assert_allclose(sum(K[3 * a:3 * a + 3, 3 * b:3 * b + 3] for b in range(4)), 0,
                                       atol=1e-9 * np.abs(K).max())
        # patch test: affine u = G X with symmetric G -> f_a = V sigma grad N_a
        Gs = rng.standard_normal((3, 3)) * 1e-3
        Gs = Gs + Gs.T
        u = X @ Gs.T
        g, V = grad_n(X)
        sig = LAM * np.trace(Gs) * np.eye(3) + 2 * MU * Gs
        np.testing.assert_allclose(K @ u.ravel(), (V * g @ sig.T).ravel(), rtol=1e-10, atol=1e-10 * np.abs(K @ u.ravel()).max())


def test_corotational_rigid_motion_and_stretch(oracle):
    rng = np.random.default_rng(2)
    for _ in range(20):
        X = tet(rng)
        Q = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        x = X @ Q.T + rng.standard_normal(3)
        K, f, R = corot(oracle, X, x)
        np.testing.assert_allclose(R, Q, atol=1e-12)
        assert np.abs(f).max() < 1e-9 * np.abs(K).max() * 0.1
        Kr = np.kron(np.eye(4), Q) @ k0(oracle, X) @ np.kron(np.eye(4), Q).T
        np.testing.assert_allclose(K, Kr, atol=1e-10 * np.abs(Kr).max())
        # symmetric stretch: R = I and f = K0 (x - X)
        Gs = rng.standard_normal((3, 3)) * 1e-2
        Gs = Gs + Gs.T
        x2 = X + X @ Gs.T
        K2, f2, R2 = corot(oracle, X, x2)
        np.testing.assert_allclose(R2, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(f2, k0(oracle, X) @ (x2 - X).ravel(), atol=1e-10 * np.abs(f2).max())
        # nodal forces balance
        np.testing.assert_allclose(f2.reshape(4, 3).sum(0), 0, atol=1e-10 * np.abs(f2).max())


def psi(F):
    J = np.linalg.det(F)
    return MU / 2 * (np.trace(F.T @ F) - 3) - MU * np.log(J) + LAM / 2 * np.log(J) ** 2


def test_neohookean_stress_matches_energy_derivative(oracle):
    rng = np.random.default_rng(3)
    for trial in range(30):
        F = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
        P = np.zeros(9); D = np.zeros(81)
        oracle.lib().oro_neo_stress(oracle._d(oracle.f64(F.ravel())), E, NU, 0.1, oracle._d(P), oracle._d(D))
        P = P.reshape(3, 3)
        h = 1e-6
        Pfd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                dF = np.zeros((3, 3)); dF[i, j] = h
                Pfd[i, j] = (psi(F + dF) - psi(F - dF)) / (2 * h)
        np.testing.assert_allclose(P, Pfd, rtol=1e-6, atol=1e-6 * np.abs(P).max())
    P = np.zeros(9)
    oracle.lib().oro_neo_stress(oracle._d(np.eye(3).ravel().copy()), E, NU, 0.1, oracle._d(P), None)
    assert np.abs(P).max() < 1e-9 * MU


def test_neohookean_element_force_and_stiffness_are_derivatives(oracle):
    rng = np.random.default_rng(4)
    checked = 0
    for trial in range(40):
        X = tet(rng)
        if trial % 2:
            Fd = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
        else:
            # mild stretch (singular values slightly > 1): alpha +- beta of every 2x2 block and the
            # 3x3 block stay positive, so the PSD projections are inactive
            Q1 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
            Q2 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
            Fd = Q1 @ np.diag(rng.uniform(1.005, 1.03, 3)) @ Q2.T
        x = X @ Fd.T
        K, f = neo(oracle, X, x)
        g, V = grad_n(X)

        def energy(xx):
            Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], 1)
            Ds = np.stack([xx[1] - xx[0], xx[2] - xx[0], xx[3] - xx[0]], 1)
            return V * psi(Ds @ np.linalg.inv(Dm))

        h = 1e-7 * np.abs(X).max()
        ffd = np.zeros(12)
        for k in range(12):
            d = np.zeros(12); d[k] = h
            ffd[k] = (energy(x + d.reshape(4, 3)) - energy(x - d.reshape(4, 3))) / (2 * h)
        np.testing.assert_allclose(f, ffd, rtol=1e-5, atol=1e-5 * np.abs(f).max())
        # stiffness = d f / d x wherever the PSD projection is inactive
        Kfd = np.zeros((12, 12))
        for k in range(12):
            d = np.zeros(12); d[k] = h
            _, fp = neo(oracle, X, x + d.reshape(4, 3))
            _, fm = neo(oracle, X, x - d.reshape(4, 3))
            Kfd[:, k] = (fp - fm) / (2 * h)
        if trial % 2 == 0:
            np.testing.assert_allclose(K, Kfd, rtol=1e-4, atol=1e-4 * np.abs(K).max())
            checked += 1
        assert np.linalg.eigvalsh(K).min() > -1e-8 * np.abs(K).max()   # projected: PSD
    assert checked >= 10


def test_mass_and_body_force_totals(oracle):
    w = W.box_small()
    s = oracle.System(w.X, w.tets, w.fixed_dofs)
    Mv = s.mass(1000.0)
    ones_x = np.tile([1.0, 0, 0], s.ndof // 3)
    total = ones_x @ s.spmdv(Mv, ones_x)
    # full-mesh mass = rho * volume when nothing is constrained
    s_free = oracle.System(w.X, w.tets, np.zeros(0, np.int32))
    Mf = s_free.mass(1000.0)
    ones_f = np.tile([1.0, 0, 0], s_free.ndof // 3)
    np.testing.assert_allclose(ones_f @ s_free.spmdv(Mf, ones_f), 1000.0 * 1.0 * 0.2 * 0.2, rtol=1e-12)
    assert total < 40.0
    fb = s_free.body_force(w.body)
    np.testing.assert_allclose(fb.reshape(-1, 3).sum(0), [0, 0, -9.81 * 40.0], rtol=1e-12)


# ---- neo-Hookean PSD projections and the inversion branch ------------------------------
# forceSemidefinite3x3 / 2x2 (element/tetelement.cpp:446-499) against numpy's eigen-clamp, the
# projected dP/dF against the block-wise clamp of a finite-difference derivative of P, and P on
# inverted or flattened tets against the signed, clamped SVD (material/invertibleHyperelastic.cpp:
# 139-245): the sign of det F goes onto the smallest singular value, which the threshold then
# clamps from below.

def eigen_clamp(A):
    lam, Q = np.linalg.eigh(A)
    return (Q * np.maximum(lam, 0.0)) @ Q.T


def test_force_semidefinite_3x3_is_the_eigen_clamp(oracle):
    rng = np.random.default_rng(21)
    projected = 0
    for trial in range(200):
        B = rng.standard_normal((3, 3))
        A = B + B.T if trial % 3 else B @ B.T              # indefinite, or PSD (left unchanged)
        packed = np.array([A[0, 0], A[0, 1], A[0, 2], A[1, 1], A[1, 2], A[2, 2]])
        out = packed.copy()
        oracle.lib().oro_psd3(oracle._d(out))
        if np.linalg.eigvalsh(A).min() >= 0:
            assert np.array_equal(out, packed)
            continue
        C = eigen_clamp(A)
        expect = np.array([C[0, 0], C[0, 1], C[0, 2], C[1, 1], C[1, 2], C[2, 2]])
        np.testing.assert_allclose(out, expect, atol=1e-12 * np.abs(A).max())
        projected += 1
    assert projected > 100


def test_force_semidefinite_2x2_is_the_eigen_clamp(oracle):
    """The block [[diag, offdiag], [offdiag, diag]] (the (ij, ji) pair of dP/dF) has eigenvalues
    diag +- offdiag; the projection clamps the negative ones to zero."""
    rng = np.random.default_rng(22)
    cases = [(d, o) for d, o in rng.standard_normal((300, 2)) * 3.0] + [(1.0, 2.0), (1.0, -2.0), (-1.0, 0.5),
                                                                         (2.0, 0.5), (-3.0, -0.5)]
    seen = set()
    for d, o in cases:
        dd, oo = C_double(d), C_double(o)
        oracle.lib().oro_psd2(oracle._d(dd), oracle._d(oo))
        A = np.array([[d, o], [o, d]])
        C = eigen_clamp(A)
        np.testing.assert_allclose([dd[0], oo[0]], [C[0, 0], C[0, 1]], atol=1e-14 * (abs(d) + abs(o)))
        seen.add(((d + o) < 0, (d - o) < 0))
    assert len(seen) == 4          # every clamp case taken


def C_double(v):
    return np.array([v], np.float64)


def neo_stress(oracle, F, thr=0.1, with_dpdf=False):
    P = np.zeros(9); T = np.zeros(81)
    oracle.lib().oro_neo_stress(oracle._d(np.ascontiguousarray(F, np.float64).ravel().copy()), E, NU, thr,
                                oracle._d(P), oracle._d(T) if with_dpdf else None)
    return (P.reshape(3, 3), T.reshape(3, 3, 3, 3)) if with_dpdf else P.reshape(3, 3)


@pytest.mark.parametrize("sigma", [(1.3, 1.1, 0.9), (0.6, 0.45, 0.3), (1.2, 0.5, 0.25), (0.95, 0.8, 0.7)])
def test_neohookean_projected_dPdF_is_the_blockwise_clamp_of_dP_dF(oracle, sigma):
    """At F = diag(sigma) (distinct, above the clamp) the oracle's projected dP/dF equals the
    finite-difference derivative of its own P (itself pinned to the closed-form energy), with
    the 3x3 block on (00, 11, 22) and the 2x2 blocks on (01, 10), (02, 20), (12, 21) each
    replaced by its eigen-clamp.  Compressed states make the 2x2 blocks indefinite."""
    F = np.diag(sigma).astype(float)
    _, T = neo_stress(oracle, F, with_dpdf=True)
    h = 1e-6
    D = np.zeros((3, 3, 3, 3))
    for k in range(3):
        for l in range(3):
            dF = np.zeros((3, 3)); dF[k, l] = h
            D[:, :, k, l] = (neo_stress(oracle, F + dF) - neo_stress(oracle, F - dF)) / (2 * h)
    expect = np.zeros((3, 3, 3, 3))
    idx3 = [(0, 0), (1, 1), (2, 2)]
    B3 = np.array([[D[i, j, k, l] for (k, l) in idx3] for (i, j) in idx3])
    C3 = eigen_clamp(0.5 * (B3 + B3.T))
    for p, (i, j) in enumerate(idx3):
        for q, (k, l) in enumerate(idx3):
            expect[i, j, k, l] = C3[p, q]
    active = 0
    for a, b in ((0, 1), (0, 2), (1, 2)):
        pair = [(a, b), (b, a)]
        B2 = np.array([[D[i, j, k, l] for (k, l) in pair] for (i, j) in pair])
        active += int(np.linalg.eigvalsh(0.5 * (B2 + B2.T)).min() < 0)
        C2 = eigen_clamp(0.5 * (B2 + B2.T))
        for p, (i, j) in enumerate(pair):
            for q, (k, l) in enumerate(pair):
                expect[i, j, k, l] = C2[p, q]
    np.testing.assert_allclose(T, expect, atol=1e-6 * np.abs(expect).max())
    if min(sigma) < 0.5:
        assert active > 0          # the compressed states exercise the 2x2 projection


def signed_clamped_svd_stress(F, thr):
    """U diag(p(s)) V^T from numpy's SVD with proper rotations U, V: when det F < 0 the smallest
    singular value carries the sign; s = max(signed sigma, thr); p_i = mu s_i - (mu - lam ln J)/s_i
    with J = s_0 s_1 s_2 (the derivative of the compressible neo-Hookean energy in the diagonal
    space)."""
    U, S, Vt = np.linalg.svd(F)
    S = S.copy()
    if np.linalg.det(U) < 0:
        U[:, 2] *= -1; S[2] *= -1
    if np.linalg.det(Vt) < 0:
        Vt[2, :] *= -1; S[2] *= -1
    s = np.maximum(S, thr)
    J = np.prod(s)
    p = MU * s - (MU - LAM * np.log(J)) / s
    return (U * p) @ Vt, S


def test_neohookean_inverted_and_flattened_stress_is_the_clamped_signed_svd(oracle):
    """Inverted, compressed-below-the-clamp and nearly flat tets in general position (the
    smallest singular value 1e-3 to 1e-2: the regular branch, then the clamp; below that the
    reference's V^T row u^T F / sigma inherits the rounding of sigma^2 and drifts from unit
    length, a conditioning property of its method, not a difference in P's formula)."""
    rng = np.random.default_rng(23)
    kinds = {"inverted": 0, "compressed": 0, "nearly flat": 0}
    for trial in range(120):
        Q1 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        Q2 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        sig = rng.uniform(0.3, 1.5, 3)
        mode = trial % 3
        if mode == 0:
            sig[rng.integers(3)] *= -rng.uniform(0.1, 1.0)     # inverted (det F < 0)
        elif mode == 1:
            sig[rng.integers(3)] = rng.uniform(0.01, 0.09)      # compressed below the clamp
        else:
            sig[rng.integers(3)] = 10.0 ** rng.uniform(-3, -2)  # nearly flat
        F = Q1 @ np.diag(sig) @ Q2.T
        P = neo_stress(oracle, F, thr=0.1)
        Pe, S = signed_clamped_svd_stress(F, 0.1)
        np.testing.assert_allclose(P, Pe, atol=1e-9 * np.abs(Pe).max())
        kinds[["inverted", "compressed", "nearly flat"][mode]] += 1
    assert min(kinds.values()) == 40


def test_neohookean_degenerate_branches_on_exactly_flat_tets(oracle):
    """F = diag(sigma) with one exact zero: the reference's closed-form eigen-solver returns the
    zero eigenvalue of F F^T as rounding noise of either sign (about 1e-16).  Below zero it is
    taken as 0 and the degenerate branch (invertibleHyperelastic.cpp:156-231) completes V^T by a
    cross product and clamps the zero singular value to the threshold: P = diag(p(s)).  Above
    zero its root (about 1.5e-8) passes the 1e-8 test, the regular branch divides the zero row
    u^T F by it, and the flat direction's term of P vanishes.  Both formulas are checked, and
    both branches must occur over the cases."""
    branches = set()
    for sigma in [(1.2, 0.7, 0.0), (0.0, 1.1, 0.9), (0.8, 0.0, 1.3), (0.0, 1.21, 0.81), (0.5, 0.0, 0.6),
                  (1.3, 0.9, 0.0), (0.0, 0.4, 0.3), (0.7, 1.7, 0.0), (2.0, 0.0, 0.05)]:
        F = np.diag(sigma).astype(float)
        P = neo_stress(oracle, F, thr=0.1)
        s = np.maximum(np.array(sigma, float), 0.1)
        J = np.prod(s)
        p = MU * s - (MU - LAM * np.log(J)) / s
        degenerate = np.diag(p)
        regular = np.diag(np.where(np.array(sigma) == 0.0, 0.0, p))
        tol = 1e-9 * np.abs(p).max()
        if np.abs(P - degenerate).max() < tol:
            branches.add("degenerate")
        else:
            np.testing.assert_allclose(P, regular, atol=tol)
            branches.add("regular")
    assert branches == {"degenerate", "regular"}


def test_neohookean_inverted_diagonal_is_reflected_and_clamped(oracle):
    """F = diag(-1.2, 0.7, 0.9): det F < 0, the smallest singular value in magnitude carries the
    sign (invertibleHyperelastic.cpp:147-152), so 0.7 becomes -0.7 and clamps to the threshold."""
    F = np.diag([-1.2, 0.7, 0.9])
    P = neo_stress(oracle, F, thr=0.1)
    Pe, S = signed_clamped_svd_stress(F, 0.1)
    np.testing.assert_allclose(P, Pe, atol=1e-9 * np.abs(Pe).max())
