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
