"""Pins the oracle's dense 3x3 numerics against numpy (SVD polar, eigh) and closed forms.

polarDecompostion3x3 (core/numerMethod.h:570-1016), properQRDecompostion3x3
(:431-507) and eigenSym3x3 (:327-429).
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation


def polar(O, F):
    F = np.ascontiguousarray(F, np.float64).ravel()
    R = np.zeros(9); S = np.zeros(9)
    O.lib().oro_polar(O._d(F), O._d(R), O._d(S))
    return R.reshape(3, 3), S.reshape(3, 3)


def svd_polar(F):
    U, s, Vt = np.linalg.svd(F)
    return U @ Vt


def sample_F(rng, kind):
    if kind == "random":
        return rng.standard_normal((3, 3))
    if kind == "near_identity":
        return np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    if kind == "rotation_stretch":
        Q = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        return Q @ np.diag(rng.uniform(0.5, 2.0, 3))
    if kind == "inverted":
        F = rng.standard_normal((3, 3))
        return F if np.linalg.det(F) < 0 else -F
    if kind == "ill_conditioned":
        Q1 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        Q2 = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
        return Q1 @ np.diag([1.0, 1e-3, 1e-6]) @ Q2
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["random", "near_identity", "rotation_stretch", "inverted", "ill_conditioned"])
def test_polar_matches_svd_polar(oracle, kind):
    rng = np.random.default_rng(hash(kind) % 2**32)
    for _ in range(200):
        F = sample_F(rng, kind)
        R, S = polar(oracle, F)
        Rs = svd_polar(F)
        tol = 1e-9 if kind == "ill_conditioned" else 1e-12
        np.testing.assert_allclose(R, Rs, atol=tol * 10)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-13)
        assert abs(np.linalg.det(R) - np.sign(np.linalg.det(F))) < 1e-12
        # S = R^T F is symmetric positive semi-definite and F = R S
        np.testing.assert_allclose(R @ S, F, atol=1e-12 * np.abs(F).max())
        np.testing.assert_allclose(S, S.T, atol=1e-9 * np.abs(F).max())
        assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > -1e-9 * np.abs(F).max()


def test_polar_special_cases(oracle):
    R, _ = polar(oracle, np.zeros((3, 3)))                       # F = 0 -> identity (numerMethod.h:576-581)
    np.testing.assert_array_equal(R, np.eye(3))
    R, _ = polar(oracle, np.diag([2.0, 0.5, 3.0]))              # diagonal positive -> identity
    np.testing.assert_allclose(R, np.eye(3), atol=1e-15)
    Q = Rotation.from_rotvec([0.3, -1.2, 2.0]).as_matrix()      # rotation -> itself
    R, _ = polar(oracle, Q)
    np.testing.assert_allclose(R, Q, atol=1e-14)


def test_qr_is_proper_rotation_times_upper(oracle):
    rng = np.random.default_rng(7)
    for _ in range(200):
        F = rng.standard_normal((3, 3)).ravel()
        Q = np.zeros(9); R = np.zeros(9)
        oracle.lib().oro_qr(oracle._d(F), oracle._d(Q), oracle._d(R))
        Q = Q.reshape(3, 3); R = R.reshape(3, 3)
        np.testing.assert_allclose(Q @ R, F.reshape(3, 3), atol=1e-13)
        np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(Q) - 1.0) < 1e-13
        assert abs(R[1, 0]) + abs(R[2, 0]) + abs(R[2, 1]) == 0.0
        assert R[0, 0] >= 0 and R[1, 1] >= 0


def test_eigsym_matches_eigh(oracle):
    rng = np.random.default_rng(11)
    for trial in range(300):
        B = rng.standard_normal((3, 3))
        A = B + B.T if trial % 3 else B @ B.T
        if trial % 7 == 0:                                       # repeated eigenvalue
            Q = Rotation.from_rotvec(rng.standard_normal(3)).as_matrix()
            A = Q @ np.diag([1.0, 1.0, 3.0]) @ Q.T
        tri = np.array([A[0, 0], A[0, 1], A[0, 2], A[1, 1], A[1, 2], A[2, 2]])
        vals = np.zeros(3); vecs = np.zeros(9)
        oracle.lib().oro_eigsym(oracle._d(tri), oracle._d(vals), oracle._d(vecs))
        V = vecs.reshape(3, 3)                                   # eigenvectors are rows
        np.testing.assert_allclose(np.sort(vals), np.linalg.eigvalsh(A), atol=1e-12 * np.abs(A).max())
        np.testing.assert_allclose(V @ V.T, np.eye(3), atol=1e-10)
        np.testing.assert_allclose(V @ A @ V.T, np.diag(vals), atol=1e-10 * np.abs(A).max())
    # scalar matrix: the nan branch returns the identity basis
    vals = np.zeros(3); vecs = np.zeros(9)
    tri = np.array([2.0, 0, 0, 2.0, 0, 2.0])
    oracle.lib().oro_eigsym(oracle._d(tri), oracle._d(vals), oracle._d(vecs))
    np.testing.assert_array_equal(vals, [2, 2, 2])
    np.testing.assert_array_equal(vecs.reshape(3, 3), np.eye(3))
