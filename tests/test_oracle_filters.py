"""Pins the oracle's EnKF, ETKF and LETKF against what the mathematics fixes:
closed forms, limits, and identities between algebraically different
formulations (Woodbury), computed here with numpy's LAPACK routines."""
import numpy as np
import pytest


def make_case(rng, k, n, p, spread=0.5):
    truth = rng.standard_normal(n)
    X = truth + spread * rng.standard_normal((k, n))
    H = np.zeros((p, n))
    for o in range(p):
        H[o, (o * 2) % n] = 0.7
        H[o, (o * 2 + 1) % n] = 0.3
    Yb = X @ H.T
    Yb = Yb.reshape(k, p)
    yo = H @ truth + rng.standard_normal(p) * 0.8
    var = rng.uniform(0.3, 1.5, p)
    Z = rng.standard_normal((p, k))
    return X, H, Yb, yo, var, Z


# ---- EnKF: EnKF.hpp:139-256 --------------------------------------------------------
def test_enkf_k2_p1_closed_form(orc):
    X = np.array([[1.0, -2.0, 0.5], [2.0, 0.0, 1.5]])
    Yb = np.array([[0.3], [1.1]])
    yo, var, Z, rho = np.array([0.9]), np.array([0.49]), np.array([[0.2, -1.3]]), 1.21
    Xa, dg = orc.enkf_analyse(X, Yb, yo, var, Z, rho)
    xm = X.mean(0); Xp = (X - xm) * np.sqrt(rho)
    yp = Yb[:, 0] - Yb[:, 0].mean()
    S = yp @ yp / 1.0 + var[0]
    K = (Xp.T @ yp) / S
    for j in range(2):
        want = xm + Xp[j] + K * (yo[0] + 0.7 * Z[0, j] - Yb[j, 0])
        np.testing.assert_allclose(Xa[j], want, rtol=1e-14, atol=1e-14)
    assert dg["condition_number"] == 1.0
    assert dg["innovation_norm"] == pytest.approx(abs(yo[0] - Yb.mean()))


def test_enkf_equals_woodbury_ensemble_space_form(orc):
    """K = X'Y'^T S^-1/(k-1) (p x p inverse, as the reference) must equal
    X' [(k-1)I + Y'^T R^-1 Y']^-1 Y'^T R^-1 (k x k inverse)."""
    rng = np.random.default_rng(11)
    k, n, p, rho = 12, 30, 17, 1.1
    X, H, Yb, yo, var, Z = make_case(rng, k, n, p)
    Xa, dg = orc.enkf_analyse(X, Yb, yo, var, Z, rho)
    xm = X.mean(0); Xp = (X - xm).T * np.sqrt(rho)          # n x k
    yb = Yb.mean(0); Yp = (Yb - yb).T                          # p x k
    Ri = 1.0 / var
    M = (k - 1) * np.eye(k) + Yp.T @ (Ri[:, None] * Yp)
    G = np.linalg.solve(M, Yp.T * Ri)                          # k x p
    D = yo[:, None] + np.sqrt(var)[:, None] * Z - Yb.T        # p x k
    want = xm[:, None] + Xp + Xp @ (G @ D)
    np.testing.assert_allclose(Xa.T, want, rtol=1e-11, atol=1e-12)
    K = Xp @ G
    assert dg["max_kalman_gain"] == pytest.approx(K.max(), rel=1e-11)
    assert dg["min_kalman_gain"] == pytest.approx(K.min(), rel=1e-11)
    S = Yp @ Yp.T / (k - 1) + np.diag(var)
    ev = np.linalg.eigvalsh(S)
    assert dg["condition_number"] == pytest.approx(ev[-1] / ev[0], rel=1e-10)
    assert dg["background_spread"] == pytest.approx(np.sqrt((Xp ** 2).mean()), rel=1e-13)
    Xan = Xa - Xa.mean(0)
    assert dg["analysis_spread"] == pytest.approx(np.sqrt((Xan ** 2).mean()), rel=1e-12)
    assert dg["analysis_increment_norm"] == 0.0


def test_enkf_limits(orc):
    rng = np.random.default_rng(12)
    k, n, p = 8, 10, 10
    X, H, Yb, yo, var, Z = make_case(rng, k, n, p)
    rho = 1.3
    Xa, _ = orc.enkf_analyse(X, Yb, yo, np.full(p, 1e30), Z * 0, rho)
    xm = X.mean(0)
    np.testing.assert_allclose(Xa, xm + (X - xm) * np.sqrt(rho), rtol=1e-12, atol=1e-12)
    # H = I, R -> 0, full-rank ensemble (k > n): members collapse onto the
    # (perturbed) observations
    k = 20
    X = rng.standard_normal((k, n)); Yb = X.copy(); yo = rng.standard_normal(n)
    Xa, _ = orc.enkf_analyse(X, Yb, yo, np.full(n, 1e-16), rng.standard_normal((n, k)), 1.0)
    np.testing.assert_allclose(Xa, np.broadcast_to(yo, Xa.shape), atol=1e-6)


# ---- ETKF: ETKF.hpp:100-178 --------------------------------------------------------
def test_etkf_against_lapack_formulation(orc):
    rng = np.random.default_rng(13)
    k, n, p, rho = 10, 25, 14, 1.07
    X, H, Yb, yo, var, Z = make_case(rng, k, n, p)
    Xa = orc.etkf_analyse(X, Yb, yo, var, rho)
    xm = X.mean(0); Xp = (X - xm).T * rho
    yb = Yb.mean(0); Yp = (Yb - yb).T; d = yo - yb
    Pa = np.linalg.inv(Yp.T @ (Yp / var[:, None]) + (k - 1) * np.eye(k))
    wa = Pa @ (Yp.T @ (d / var))
    Wa = np.sqrt(k - 1) * np.linalg.cholesky(Pa)
    np.testing.assert_allclose(Wa @ Wa.T, (k - 1) * Pa, rtol=1e-12)
    want = (xm + Xp @ wa)[:, None] + Xp @ Wa
    np.testing.assert_allclose(Xa.T, want, rtol=1e-11, atol=1e-12)


# ---- LETKF ---------------------------------------------------------------------
def grid_problem(orc, rng, nx, ny, nlev, k, p, radius, rho=1.05, coord_obs=None):
    n = nx * ny * nlev
    X = rng.standard_normal((k, n))
    gx, gy = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float))
    col_a, col_b = gx.ravel(), gy.ravel()
    if coord_obs is None:
        ox, oy = rng.uniform(0, nx - 1, p), rng.uniform(0, ny - 1, p)
    else:
        ox, oy = coord_obs
    lev = rng.integers(0, nlev, p)
    idx = (lev * nx * ny + np.rint(oy).astype(int) * nx + np.rint(ox).astype(int))
    Yb = X[:, idx] + 0.1 * X[:, (idx + 1) % n]
    yo = rng.standard_normal(len(ox))
    var = rng.uniform(0.2, 0.6, len(ox))
    prob = orc.Problem(k, nx, ny, nlev, orc.COORD_GRID, col_a, col_b, ox, oy, yo, var,
                       rho, radius)
    return prob, X, Yb, yo, var


def test_letkf_standard_infinite_radius_is_global_hunt_etkf(orc):
    rng = np.random.default_rng(21)
    k, rho = 9, 1.2
    prob, X, Yb, yo, var = grid_problem(orc, rng, 5, 4, 3, k, 30, 1e9, rho)
    Xa, T, nloc = orc.letkf_standard(prob, X, Yb, want_T=True)
    assert np.all(nloc == 30)
    yb = Yb.mean(0); Yp = (Yb - yb).T; d = yo - yb
    A = (k - 1) / rho * np.eye(k) + Yp.T @ (Yp / var[:, None])  # GC(0) = 1
    w, V = np.linalg.eigh(A)
    wbar = V @ ((V.T @ (Yp.T @ (d / var))) / w)
    W = np.sqrt(k - 1) * (V / np.sqrt(w)) @ V.T
    Tw = wbar[:, None] + W
    np.testing.assert_allclose(T, np.broadcast_to(Tw, T.shape), rtol=1e-11, atol=1e-13)
    xm = X.mean(0)
    np.testing.assert_allclose(Xa, (xm[:, None] + (X - xm).T @ Tw).T, rtol=1e-12, atol=1e-12)


def test_letkf_standard_transform_invariants(orc):
    rng = np.random.default_rng(22)
    k, rho = 16, 1.05
    prob, X, Yb, yo, var = grid_problem(orc, rng, 8, 6, 2, k, 60, 3.0, rho)
    Xa, T, nloc = orc.letkf_standard(prob, X, Yb, want_T=True)
    for g in range(T.shape[0]):
        # T = wbar 1^T + W with W symmetric and W 1 = sqrt(rho) 1 (1 is an
        # eigenvector of A with eigenvalue (k-1)/rho because Y' 1 = 0)
        wbar = (T[g] @ np.ones(k) - np.sqrt(rho)) / k
        Wg = T[g] - wbar[:, None]
        np.testing.assert_allclose(Wg, Wg.T, atol=1e-13)
        np.testing.assert_allclose(Wg @ np.ones(k), np.sqrt(rho), rtol=1e-12)
    # analysis perturbations stay zero-mean: mean(Xa) = xbar + X' wbar
    xm = X.mean(0)
    Xam = Xa.mean(0)
    ncol = 48
    for g in range(ncol):
        wbar = (T[g] @ np.ones(k) - np.sqrt(rho)) / k
        for l in range(2):
            i = l * ncol + g
            assert Xam[i] == pytest.approx(xm[i] + (X[:, i] - xm[i]) @ wbar, abs=1e-12)


def test_letkf_empty_column_closed_form(orc):
    rng = np.random.default_rng(23)
    k, rho = 7, 1.3
    prob, X, Yb, yo, var = grid_problem(orc, rng, 6, 5, 2, k, 1, 0.5, rho,
                                        coord_obs=(np.array([0.0]), np.array([0.0])))
    Xa, T, nloc = orc.letkf_standard(prob, X, Yb, want_T=True)
    xm = X.mean(0)
    want = xm + np.sqrt(rho) * (X - xm)
    # every column except (0,0) has no local obs
    for g in range(1, 30):
        assert nloc[g] == 0
        for l in range(2):
            i = l * 30 + g
            np.testing.assert_allclose(Xa[:, i], want[:, i], rtol=1e-14, atol=1e-14)


def test_letkf_refcompat_closed_form_k2(orc):
    """One local obs, k = 2: Q2 (R = I), Q3 (Pa *= rho after the inverse),
    Q4 (Cholesky square root) and Q5 (element-wise combination)."""
    rho = 1.4
    X = np.array([[0.3], [1.7]])
    Yb = np.array([[0.2], [1.0]])
    yo = np.array([0.5])
    prob = orc.Problem(2, 1, 1, 1, orc.COORD_GRID, [0.0], [0.0], [0.0], [0.0], yo,
                       np.array([0.25]), rho, 1.0)
    Xa, t = orc.letkf_refcompat_snapshot(prob, X, Yb, want_t=True)
    a = (Yb[1, 0] - Yb[0, 0]) / 2           # Y' = [-a, a]
    d = yo[0] - Yb.mean()
    Pa = rho / (2 * a * a + 1) * np.array([[a * a + 1, a * a], [a * a, a * a + 1]])
    wa = rho / (2 * a * a + 1) * np.array([-a, a]) * d
    L11 = np.sqrt(Pa[0, 0]); L21 = Pa[1, 0] / L11; L22 = np.sqrt(Pa[1, 1] - L21 ** 2)
    rs = np.array([L11, L21 + L22])          # sqrt(k-1) = 1
    xm = X.mean(); xp = X[:, 0] - xm
    np.testing.assert_allclose(Xa[:, 0], xm + xp * wa + xp * rs, rtol=1e-14)
    np.testing.assert_allclose(t[0], wa + rs, rtol=1e-14)


def test_letkf_inplace_equals_snapshot_when_points_decouple(orc):
    rng = np.random.default_rng(24)
    k, nx, ny = 6, 5, 4
    n = nx * ny
    X = rng.standard_normal((k, n))
    gx, gy = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float))
    ox, oy = gx.ravel()[::2].copy(), gy.ravel()[::2].copy()
    idx = (oy * nx + ox).astype(np.int64)
    rp = np.arange(len(idx) + 1, dtype=np.int64)
    yo = rng.standard_normal(len(idx))
    prob = orc.Problem(k, nx, ny, 1, orc.COORD_GRID, gx.ravel(), gy.ravel(), ox, oy, yo,
                       np.ones(len(idx)), 1.1, 0.5)
    Yb = X[:, idx]
    a = orc.letkf_refcompat_snapshot(prob, X, Yb)
    b = orc.letkf_refcompat_inplace(prob, X, rp, idx, np.ones(len(idx)))
    np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14)
    # with overlapping local sets the literal in-place order differs (Q1)
    prob2 = orc.Problem(k, nx, ny, 1, orc.COORD_GRID, gx.ravel(), gy.ravel(), ox, oy, yo,
                        np.ones(len(idx)), 1.1, 3.0)
    a2 = orc.letkf_refcompat_snapshot(prob2, X, Yb)
    b2 = orc.letkf_refcompat_inplace(prob2, X, rp, idx, np.ones(len(idx)))
    assert np.abs(a2 - b2).max() > 1e-6


def test_letkf_geographic_band_index_matches_full_scan(orc):
    """The latitude-band index must find exactly the obs a full scan finds."""
    rng = np.random.default_rng(25)
    k, nx, ny = 5, 16, 8
    lon = (np.arange(nx) + 0.5) * 360 / nx
    lat = -90 + (np.arange(ny) + 0.5) * 180 / ny
    LON, LAT = np.meshgrid(lon, lat)
    p = 400
    olat = np.degrees(np.arcsin(rng.uniform(-1, 1, p)))
    olon = rng.uniform(0, 360, p)
    X = rng.standard_normal((k, nx * ny))
    Yb = rng.standard_normal((k, p))
    yo = rng.standard_normal(p)
    r = 2500.0
    prob = orc.Problem(k, nx, ny, 1, orc.COORD_GEOGRAPHIC, LAT.ravel(), LON.ravel(), olat, olon,
                       yo, np.full(p, 0.25), 1.05, r)
    _, _, nloc = orc.letkf_standard(prob, X, Yb)
    for g in range(nx * ny):
        dd = [orc.haversine_km(LAT.ravel()[g], LON.ravel()[g], olat[o], olon[o]) for o in range(p)]
        assert nloc[g] == sum(1 for v in dd if v <= r)
