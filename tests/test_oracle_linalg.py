"""Pins the oracle's small dense linear algebra against LAPACK fixtures
(tests/golden/linalg_fixtures.npz, written by tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "linalg_fixtures.npz"))
SIZES = (2, 8, 20, 64)


@pytest.mark.parametrize("n", SIZES)
def test_cholesky_matches_lapack(orc, n):
    L = orc.cholesky_lower(FIX[f"A{n}"])
    np.testing.assert_allclose(L, FIX[f"chol{n}"], rtol=1e-10, atol=1e-12)
    assert np.allclose(np.triu(L, 1), 0.0)


@pytest.mark.parametrize("n", SIZES)
def test_lu_inverse_matches_lapack(orc, n):
    for key, ref in ((f"A{n}", f"inv{n}"), (f"M{n}", f"Minv{n}")):
        Ai = orc.lu_inverse(FIX[key])
        np.testing.assert_allclose(Ai, FIX[ref], rtol=1e-9, atol=1e-12 * np.abs(FIX[ref]).max())


def test_lu_inverse_needs_pivoting(orc):
    # zero leading pivot: only a pivoting LU inverts this
    A = np.array([[0.0, 2.0, 1.0], [1.0, 1.0, 0.0], [3.0, 0.0, 1.0]])
    np.testing.assert_allclose(orc.lu_inverse(A) @ A, np.eye(3), atol=1e-14)


@pytest.mark.parametrize("n", SIZES)
def test_jacobi_eigh_matches_lapack(orc, n):
    A = FIX[f"A{n}"]
    w, V, sweeps = orc.jacobi_eigh(A)
    np.testing.assert_allclose(w, FIX[f"eigvals{n}"], rtol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-13)
    np.testing.assert_allclose((V * w) @ V.T, A, rtol=0, atol=1e-12 * np.abs(A).max())
    S = (V / np.sqrt(w)) @ V.T
    np.testing.assert_allclose(S, FIX[f"invsqrt{n}"], rtol=1e-11, atol=1e-14)
    assert sweeps < 20


def test_jacobi_ensemble_shaped_matrix(orc):
    w, V, _ = orc.jacobi_eigh(FIX["Aens"])
    np.testing.assert_allclose(w, FIX["eigvals_ens"], rtol=1e-12)
    S = (V / np.sqrt(w)) @ V.T
    np.testing.assert_allclose(S, FIX["invsqrt_ens"], rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("n", SIZES)
def test_singular_values_match_lapack(orc, n):
    s = orc.singular_values(FIX[f"G{n}"])
    np.testing.assert_allclose(s, FIX[f"sv{n}"], rtol=1e-11)


def test_cholesky_rejects_indefinite(orc):
    with pytest.raises(RuntimeError):
        orc.cholesky_lower(np.array([[1.0, 2.0], [2.0, 1.0]]))
