"""Known-answer tests for the oracle's scalar pieces: the reference's own
MetricsTest values, haversine and Gaspari-Cohn closed forms, the IDW4 stencil
and the Lorenz-96 model."""
import math

import numpy as np
import pytest

# ---- Metrics: tests/framework/algorithms/MetricsTest.cpp:24-195 -----------------
TRUTH = np.array([1.0, 2.0, 3.0, 4.0])
ENS = np.array([[1.1, 2.1, 3.1, 4.1], [0.9, 1.9, 2.9, 3.9], [1.0, 2.0, 3.0, 4.0]])


def test_metrics_reference_kats(orc):
    m = orc.metrics(ENS, TRUTH)
    np.testing.assert_allclose(m["mean"], TRUTH, atol=1e-6)
    np.testing.assert_allclose(m["spread"], 0.1, atol=1e-6)
    assert abs(m["bias"]) < 1e-6
    assert abs(m["correlation"] - 1.0) < 1e-6
    assert abs(m["crps"] - 0.022222222222222227) < 1e-6 and m["crps"] > 0
    assert abs(m["rmse"]) < 1e-6
    assert abs(m["avg_spread"] - 0.1) < 1e-6


def test_metrics_biased_ensemble(orc):
    m = orc.metrics(ENS + 0.5, TRUTH)
    assert abs(m["bias"] - 0.5) < 1e-6
    assert abs(m["rmse"] - 0.5) < 1e-6
    assert abs(m["correlation"] - 1.0) < 1e-6


def test_metrics_uncorrelated_ensemble(orc):
    ens = np.array([[10.0, 2.0, 300.0, 4.0], [4.0, 3.0, 20.0, 1.0], [2.0, 30.0, 1.0, 43.0]])
    m = orc.metrics(ens, TRUTH)
    assert abs(m["correlation"] - 0.34105511573854358) < 1e-6
    assert m["rmse"] > 1.0 and m["crps"] > 1.0


# ---- haversine: Location.hpp:349-357 -----------------------------------------------
def test_haversine_closed_forms(orc):
    R = 6371.0
    assert abs(orc.haversine_km(0, 0, 90, 0) - math.pi * R / 2) < 1e-9
    assert abs(orc.haversine_km(0, 0, 0, 180) - math.pi * R) < 1e-9
    assert abs(orc.haversine_km(0, 10, 0, 11) - R * math.radians(1)) < 1e-9
    assert orc.haversine_km(12.5, -40.0, 12.5, -40.0) == 0.0
    a = orc.haversine_km(33.0, 100.0, -12.0, 250.0)
    b = orc.haversine_km(-12.0, 250.0, 33.0, 100.0)
    assert abs(a - b) < 1e-9


def test_grid_distance(orc):
    assert orc.grid_distance(0, 0, 3, 4) == 5.0


# ---- Gaspari & Cohn (1999) eq. 4.10 ----------------------------------------------------
def test_gaspari_cohn_values(orc):
    gc = orc.gaspari_cohn
    assert gc(0.0) == 1.0
    assert gc(2.0) == 0.0 and gc(2.5) == 0.0
    assert abs(gc(1.0) - 5.0 / 24.0) < 1e-15
    e = 1e-7  # continuity of value and slope across z = 1 and z = 2
    assert abs(gc(1 - e) - gc(1 + e)) < 1e-6
    s_lo = (gc(1 - e) - gc(1 - 2 * e)) / e
    s_hi = (gc(1 + 2 * e) - gc(1 + e)) / e
    assert abs(s_lo - s_hi) < 1e-5
    assert abs(gc(2 - e)) < 1e-12
    zs = np.linspace(0, 2, 401)
    vals = np.array([gc(z) for z in zs])
    assert np.all(np.diff(vals) <= 1e-15) and np.all(vals >= 0)


def test_gaspari_cohn_is_positive_definite_in_3d(orc):
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 3, size=(120, 3))
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    C = np.vectorize(orc.gaspari_cohn)(d / 1.0)
    assert np.linalg.eigvalsh(C).min() > -1e-10


# ---- IDW4 observation operator: IdentityObsOperator.hpp:594-676 -------------------
def test_idw4_exact_hit_on_one_row_grid(orc):
    nx = 40
    x = np.linspace(-3, 5, nx) ** 2
    for i in (0, 7, 38):
        want = (2e12 * x[i] + 2 * x[i + 1]) / (2e12 + 2)
        assert orc.idw4(x, nx, 1, 1, float(i), 0.0, 0.0) == pytest.approx(want, rel=1e-15)
    assert orc.idw4(x, nx, 1, 1, float(nx - 1), 0.0, 0.0) == x[nx - 1]


def test_idw4_brute_force_2d_and_3d(orc):
    rng = np.random.default_rng(5)
    nx, ny, nz = 7, 5, 4
    st = rng.standard_normal(nx * ny * nz)
    for _ in range(50):
        x, y, z = rng.uniform(-1, nx), rng.uniform(-1, ny), rng.uniform(-1, nz)
        # literal restatement of the reference stencil, independent code
        cx, cy, cz = min(max(x, 0), nx - 1), min(max(y, 0), ny - 1), min(max(z, 0), nz - 1)
        i0, j0, k0 = int(math.floor(cx)), int(math.floor(cy)), int(math.floor(cz))
        i1, j1, k1 = min(i0 + 1, nx - 1), min(j0 + 1, ny - 1), min(k0 + 1, nz - 1)
        corners = [(i0, j0, k0), (i1, j0, k0), (i0, j1, k0), (i1, j1, k0),
                   (i0, j0, k1), (i1, j0, k1), (i0, j1, k1), (i1, j1, k1)]
        ds = sorted(((math.dist((cx, cy, cz), c), n_, c) for n_, c in enumerate(corners)))[:4]
        ws = [1e12 if d == 0 else 1 / d for d, _, _ in ds]
        val = sum(w * st[c[2] * nx * ny + c[1] * nx + c[0]] for w, (_, _, c) in zip(ws, ds))
        val /= sum(ws)
        assert orc.idw4(st, nx, ny, nz, x, y, z) == pytest.approx(val, rel=1e-12, abs=1e-12)


def test_idw4_stencil_as_csr_matches_literal(orc):
    rng = np.random.default_rng(6)
    nx, ny, nz = 9, 6, 3
    st = rng.standard_normal(nx * ny * nz)
    pts = rng.uniform(0, 1, (30, 3)) * [nx - 1, ny - 1, nz - 1]
    rp, cols, ws = [0], [], []
    for x, y, z in pts:
        idx, w = orc.idw4_stencil(nx, ny, nz, x, y, z)
        cols += list(idx); ws += list(w / w.sum()); rp.append(len(cols))
    hx = orc.csr_apply(np.array(rp), np.array(cols), np.array(ws), st)
    lit = [orc.idw4(st, nx, ny, nz, *pt) for pt in pts]
    np.testing.assert_allclose(hx, lit, rtol=1e-13, atol=1e-13)


# ---- Lorenz-96: tests/python/lorenz95.py:3-37 ----------------------------------------
def test_l96_stationary_point(orc):
    x = np.full(40, 8.0)
    assert np.all(orc.l96_tendency(x, 8.0) == 0.0)
    np.testing.assert_array_equal(orc.l96_step(x, 8.0, 0.05, "rk4"), x)


def test_l96_tendency_cyclic_indexing(orc):
    x = np.arange(6, dtype=float)
    d = orc.l96_tendency(x, 8.0)
    for i in range(6):
        want = (x[(i + 1) % 6] - x[i - 2]) * x[i - 1] - x[i] + 8.0
        assert d[i] == want


def test_l96_euler_matches_reference_step(orc):
    rng = np.random.default_rng(1)
    x = 8 + rng.standard_normal(40)
    d = orc.l96_tendency(x, 8.0)
    np.testing.assert_array_equal(orc.l96_step(x, 8.0, 0.01, "euler"), x + 0.01 * d)


def test_l96_rk4_is_fourth_order(orc):
    rng = np.random.default_rng(2)
    x0 = 8 + rng.standard_normal(40)

    def run(dt, T=0.4):
        x = x0.copy()
        for _ in range(int(round(T / dt))):
            x = orc.l96_step(x, 8.0, dt, "rk4")
        return x
    ref = run(0.4 / 1024)
    e1 = np.abs(run(0.02) - ref).max()
    e2 = np.abs(run(0.01) - ref).max()
    assert 12 < e1 / e2 < 20


def test_l96_advection_conserves_energy(orc):
    """The advection term a_i = (x_{i+1} - x_{i-2}) x_{i-1} conserves
    sum x_i^2: sum_i x_i a_i = 0 for any x (each triple product appears once
    with each sign under the cyclic shift).  A wrong index offset in the
    tendency breaks the identity."""
    rng = np.random.default_rng(12)
    for N in (5, 6, 40, 97):
        x = rng.standard_normal(N) * 3
        F = 8.0
        adv = orc.l96_tendency(x, F) + x - F
        assert abs(float(np.dot(x, adv))) < 1e-12 * float(np.dot(x, x)) * np.abs(x).max()


def test_l96_tendency_hand_computed(orc):
    # x = (1, 2, 3, 4, 5), F = 8, worked by hand:
    # i=0: (x1 - x3) x4 - x0 + F = (2-4)*5 - 1 + 8 = -3
    # i=1: (x2 - x4) x0 - x1 + F = (3-5)*1 - 2 + 8 =  4
    # i=2: (x3 - x0) x1 - x2 + F = (4-1)*2 - 3 + 8 = 11
    # i=3: (x4 - x1) x2 - x3 + F = (5-2)*3 - 4 + 8 = 13
    # i=4: (x0 - x2) x3 - x4 + F = (1-3)*4 - 5 + 8 = -5
    d = orc.l96_tendency(np.array([1.0, 2.0, 3.0, 4.0, 5.0]), 8.0)
    np.testing.assert_array_equal(d, [-3.0, 4.0, 11.0, 13.0, -5.0])


def test_cart_distance_known_answers(orc):
    assert orc.cart_distance(0, 0, 0, 1, 2, 2) == 3.0
    assert orc.cart_distance(1, 1, 1, 4, 5, 13) == 13.0       # (3, 4, 12)
    assert orc.cart_distance(-2, 5, 7, -2, 5, 7) == 0.0
    assert orc.cart_distance(1, 2, 3, 4, 6, 3) == orc.cart_distance(4, 6, 3, 1, 2, 3) == 5.0


def _one_column_letkf(orc, dist, radius, k=6, seed=0, sigma=0.7, rho=1.05, nlev=3):
    """A 1 x 1 GRID column at (0, 0) with one observation at (dist, 0)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((k, nlev))
    Yb = rng.standard_normal((k, 1))
    yo = np.array([0.8])
    prob = orc.Problem(k, 1, 1, nlev, orc.COORD_GRID, np.zeros(1), np.zeros(1),
                       np.array([dist]), np.zeros(1), yo, np.array([sigma * sigma]), rho, radius)
    Xa, _, nloc = orc.letkf_standard(prob, X, Yb)
    return X, Yb, yo, Xa, int(nloc[0])


def _single_obs_mean_increment(X, Yb, yo, w, rho):
    """Closed form of the standard LETKF mean update with one observation of
    weight w (GC taper / variance): A = (k-1)/rho I + w y' y'^T, so by
    Sherman-Morrison A^-1 y' = y' / ((k-1)/rho + w |y'|^2) and
    xbar_a - xbar = X'^T A^-1 y' w d."""
    k = X.shape[0]
    yp = Yb[:, 0] - Yb[:, 0].mean()
    d = yo[0] - Yb[:, 0].mean()
    Xp = X - X.mean(0)
    return Xp.T @ yp * (w * d / ((k - 1) / rho + w * (yp @ yp)))


def test_letkf_taper_scale_is_half_the_radius(orc):
    """GC support = radius, i.e. c = r/2: an observation at d = r/2 has
    weight GC(1) / sigma^2 = (5/24) / sigma^2 (with c = r it would be
    GC(1/2) / sigma^2 = 0.6823 / sigma^2), one at d = r has weight 0 but is
    local, one beyond r is not local."""
    r, sigma, rho = 10.0, 0.7, 1.05
    X, Yb, yo, Xa, nloc = _one_column_letkf(orc, r / 2, r, sigma=sigma, rho=rho)
    assert nloc == 1
    inc = Xa.mean(0) - X.mean(0)
    want = _single_obs_mean_increment(X, Yb, yo, (5.0 / 24.0) / sigma ** 2, rho)
    np.testing.assert_allclose(inc, want, rtol=1e-12, atol=1e-14)
    wrong = _single_obs_mean_increment(X, Yb, yo, orc.gaspari_cohn(0.5) / sigma ** 2, rho)
    assert np.max(np.abs(inc - wrong)) > 1e-3 * np.max(np.abs(want))
    # d = r: zero weight, so only the inflation acts; still a local observation
    X, Yb, yo, Xa, nloc = _one_column_letkf(orc, r, r, sigma=sigma, rho=rho)
    assert nloc == 1
    np.testing.assert_allclose(Xa.mean(0), X.mean(0), rtol=0, atol=1e-13)
    Xp, Xap = X - X.mean(0), Xa - Xa.mean(0)
    np.testing.assert_allclose(Xap, np.sqrt(rho) * Xp, rtol=1e-12, atol=1e-14)
    # beyond r: an empty column (LETKF.hpp:167-190)
    X, Yb, yo, Xa, nloc = _one_column_letkf(orc, np.nextafter(r, 20.0), r)
    assert nloc == 0
