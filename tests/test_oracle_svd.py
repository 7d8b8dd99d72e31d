"""Pins the oracle's signed 2x2 SVD (helper_math.cuh:26-140, givens.cuh:30-176;
wrapper gpu_implementation5.cu:399-416) against LAPACK and closed forms."""
import numpy as np
import pytest

import oracle


def rot(c, s):
    return np.array([[c, s], [-s, c]])


def check_decomp(F, U, S, V, tol=4e-15):
    F = np.asarray(F, dtype=np.float64)
    scale = max(np.abs(F).max(), 1e-300)
    # U, V are rotations in the reference's [[c, s], [-s, c]] form
    for R in (U, V):
        assert R[0, 0] == R[1, 1] and R[0, 1] == -R[1, 0]
        assert abs(np.linalg.det(R) - 1) < 1e-14
    rec = U @ np.diag(S) @ V.T
    assert np.abs(rec - F).max() <= tol * scale * 4
    # sorted descending, sign only on the smaller-magnitude value
    assert S[0] >= S[1]
    assert abs(S[0]) >= abs(S[1]) * (1 - 1e-15)


def test_random_against_lapack():
    rng = np.random.default_rng(1)
    for _ in range(3000):
        F = rng.normal(size=(2, 2))
        if rng.random() < 0.3:
            F = np.eye(2) + 1e-3 * F          # near identity, as in the solver
        U, S, V = oracle.svd2x2(F)
        check_decomp(F, U, S, V)
        ref = np.linalg.svd(F, compute_uv=False)
        got = np.sort(np.abs(S))[::-1]
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-15)
        # the signed product is the determinant
        assert abs(S[0] * S[1] - np.linalg.det(F)) <= 1e-14 * max(1, abs(np.linalg.det(F)))


@pytest.mark.parametrize("F,S", [
    (np.eye(2), (1.0, 1.0)),
    (np.diag([2.0, 0.5]), (2.0, 0.5)),
    (np.diag([0.5, 2.0]), (2.0, 0.5)),
    (np.array([[0.0, 1.0], [1.0, 0.0]]), (1.0, -1.0)),
    (np.diag([1.0, -1.0]), (1.0, -1.0)),
    (np.diag([3.0, 3.0]), (3.0, 3.0)),
])
def test_special_cases(F, S):
    U, Sg, V = oracle.svd2x2(F)
    assert np.allclose(Sg, S, rtol=0, atol=1e-15)
    check_decomp(F, U, Sg, V)


def test_zero_matrix_takes_identity_rotations():
    U, S, V = oracle.svd2x2(np.zeros((2, 2)))
    assert np.array_equal(S, [0.0, 0.0])
    assert np.array_equal(U, np.eye(2)) and np.array_equal(V, np.eye(2))


def test_pure_rotation_has_unit_singular_values():
    for th in np.linspace(-3, 3, 25):
        R = rot(np.cos(th), np.sin(th))
        U, S, V = oracle.svd2x2(R)
        assert np.allclose(S, [1, 1], atol=1e-15)
        assert np.allclose(U @ V.T, R, atol=1e-15)


def test_transposed_input_is_distinguished():
    # F and F^T share singular values but not U/V: the row-major input
    # convention must be kept
    F = np.array([[1.2, 0.4], [-0.1, 0.9]])
    U, S, V = oracle.svd2x2(F)
    Ut, St, Vt = oracle.svd2x2(F.T)
    assert np.allclose(S, St, atol=1e-15)
    assert np.abs(U @ np.diag(S) @ V.T - F).max() < 1e-15
    assert np.abs(Ut @ np.diag(St) @ Vt.T - F.T).max() < 1e-15
    assert not np.allclose(U, Ut)
