"""Writes the committed linear-algebra fixtures used to pin the oracle.

The expected outputs come from numpy.linalg / scipy.linalg (LAPACK), an
implementation independent of the oracle's own Jacobi, LU and Cholesky code.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))


def spd(rng, n, cond_scale):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.exp(rng.uniform(0, np.log(cond_scale), n))
    return (Q * ev) @ Q.T


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for n in (2, 8, 20, 64):
        A = spd(rng, n, 1e3)
        A = 0.5 * (A + A.T)
        w, V = np.linalg.eigh(A)
        out[f"A{n}"] = A
        out[f"eigvals{n}"] = w
        out[f"chol{n}"] = sla.cholesky(A, lower=True)
        out[f"inv{n}"] = np.linalg.inv(A)
        out[f"invsqrt{n}"] = (V / np.sqrt(w)) @ V.T
        G = rng.standard_normal((n + 3, n))
        out[f"G{n}"] = G
        out[f"sv{n}"] = np.linalg.svd(G, compute_uv=False)
        M = rng.standard_normal((n, n)) + n * np.eye(n)
        out[f"M{n}"] = M
        out[f"Minv{n}"] = np.linalg.inv(M)
    # an ensemble-shaped A = (k-1)/rho I + Y'^T R^-1 Y' with a rank-deficient Y'
    k, p = 64, 40
    Yp = rng.standard_normal((p, k)); Yp -= Yp.mean(axis=1, keepdims=True)
    A = (k - 1) / 1.05 * np.eye(k) + Yp.T @ Yp / 0.25
    w, V = np.linalg.eigh(A)
    out["Aens"] = A
    out["eigvals_ens"] = w
    out["invsqrt_ens"] = (V / np.sqrt(w)) @ V.T
    np.savez_compressed(os.path.join(HERE, "linalg_fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "linalg_fixtures.npz"))


if __name__ == "__main__":
    main()
