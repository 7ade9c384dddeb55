"""GPU tests of input validation, failure paths and coordinate systems:
CARTESIAN localisation, GEOGRAPHIC local sets at the exact radius boundary,
invalid (+inf variance) observations, non-finite inputs, k < 2, p = 0,
non-SPD factorisations on a device-resident ensemble, and the EnKF beyond
k = 64 and p = 128 -- each against the oracle or a closed form."""
import math

import numpy as np
import pytest

from cases import l96_case, latlon_case, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def mda():
    import metada_amd
    return metada_amd


@pytest.fixture(scope="module")
def ctx(mda):
    c = mda.Context(0)
    yield c
    c.close()


# ---- CARTESIAN localisation (Location.hpp:217-224) ----------------------------------------
@pytest.mark.parametrize("form", ["standard", "refcompat"])
def test_letkf_cartesian_matches_oracle(mda, ctx, orc, form):
    rng = np.random.default_rng(21)
    nx, ny, nlev, k, p = 20, 14, 3, 10, 260
    X = rng.standard_normal((k, nx * ny * nlev))
    cx = np.tile(np.arange(nx, dtype=float), ny) * 1.5
    cy = np.repeat(np.arange(ny, dtype=float), nx) * 1.5
    cz = rng.uniform(-2, 2, nx * ny)
    ox, oy, oz = rng.uniform(0, 30, p), rng.uniform(0, 21, p), rng.uniform(-2, 2, p)
    yo, var = rng.standard_normal(p), rng.uniform(0.3, 1.5, p)
    Yb = rng.standard_normal((k, p))
    obs = mda.Observation(yo, var, ox, oy, oz, coord=mda._lib.COORD_CARTESIAN)
    geom = mda.ColumnGeometry(mda._lib.COORD_CARTESIAN, cx, cy, cz)
    ens = mda.Ensemble(X.copy(), nx, ny, nlev)
    f = mda.LETKF(ens, obs, mda.ObsOperator(hx=Yb), mda.Config({"inflation": 1.05,
                                                                "localization_radius": 4.5}),
                  geometry=geom, ctx=ctx, form=form)
    f.Analyse()
    prob = orc.Problem(k, nx, ny, nlev, orc.COORD_CARTESIAN, cx, cy, ox, oy, yo, var,
                       f.inflation, f.localization_radius, col_c=cz, obs_c=oz)
    if form == "standard":
        Xo, _, nloc = orc.letkf_standard(prob, X, Yb)
        np.testing.assert_array_equal(f.local_counts(), nloc)
    else:
        Xo = orc.letkf_refcompat_snapshot(prob, X, Yb)
    assert f.local_counts().max() > 3
    assert rel_err(ens.data, Xo) < TOL


# ---- GEOGRAPHIC local sets at the radius boundary ---------------------------------------------
def _boundary_lat(orc, lat0, lon0, r, sign):
    """The last double latitude along the meridian through (lat0, lon0),
    going north (sign +1) or south (-1), whose haversine distance (the
    oracle's, i.e. the reference formula with the C library's sin/cos/atan2)
    is <= r."""
    step = math.degrees(r / 6371.0)
    lo, hi = lat0, lat0 + sign * 1.5 * step  # hav(lo) <= r < hav(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if orc.haversine_km(lat0, lon0, mid, lon0) <= r:
            lo = mid
        else:
            hi = mid
    x = lo
    toward = math.copysign(math.inf, sign)
    while orc.haversine_km(lat0, lon0, np.nextafter(x, toward), lon0) <= r:
        x = float(np.nextafter(x, toward))
    while orc.haversine_km(lat0, lon0, x, lon0) > r:
        x = float(np.nextafter(x, -toward))
    return x


def _boundary_lon(orc, lat0, lon0, r):
    """The last double longitude east of lon0 at latitude lat0 within r."""
    lo, hi = lon0, lon0 + 1.5 * math.degrees(r / 6371.0) / math.cos(math.radians(lat0))
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if orc.haversine_km(lat0, lon0, lat0, mid) <= r:
            lo = mid
        else:
            hi = mid
    x = lo
    while orc.haversine_km(lat0, lon0, lat0, np.nextafter(x, math.inf)) <= r:
        x = float(np.nextafter(x, math.inf))
    while orc.haversine_km(lat0, lon0, lat0, x) > r:
        x = float(np.nextafter(x, -math.inf))
    return x


@pytest.mark.parametrize("radius", [1500.0, 37.0, 5.0])
def test_geographic_boundary_local_sets_match_oracle(mda, ctx, orc, radius):
    """Observations at the radius and one or two ulps either side, along a
    meridian and along a parallel of several columns: every column's local
    count equals the oracle's exactly, in the FP64 and the reduced-precision
    Gram and in the hard-cutoff (reference) form, whose analysis then matches
    the oracle.  5 km is below the bucket grid's 1/16-degree floor."""
    nlon, nlat, nlev, k = 24, 12, 2, 8
    lonc = (np.arange(nlon) + 0.5) * 360 / nlon
    latc = -90 + (np.arange(nlat) + 0.5) * 180 / nlat
    LON, LAT = np.meshgrid(lonc, latc)
    col_lat, col_lon = LAT.ravel().copy(), LON.ravel().copy()
    lat, lon = [], []
    for g in (30, 77, 100, 155, 190, 250):
        la0, lo0 = col_lat[g], col_lon[g]
        for sign in (1, -1):
            b = _boundary_lat(orc, la0, lo0, radius, sign)
            toward = math.copysign(math.inf, sign)
            pts = [np.nextafter(b, -toward), b, np.nextafter(b, toward),
                   np.nextafter(np.nextafter(b, toward), toward)]
            lat += [float(v) for v in pts]
            lon += [lo0] * 4
        b = _boundary_lon(orc, la0, lo0, radius)
        pts = [np.nextafter(b, -math.inf), b, np.nextafter(b, math.inf),
               np.nextafter(np.nextafter(b, math.inf), math.inf)]
        lat += [la0] * 4
        lon += [float(v) for v in pts]
    lat, lon = np.array(lat), np.mod(np.array(lon), 360.0)
    p = len(lat)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((k, nlon * nlat * nlev))
    Yb = rng.standard_normal((k, p))
    yo, var = rng.standard_normal(p), np.full(p, 0.5)
    prob = orc.Problem(k, nlon, nlat, nlev, orc.COORD_GEOGRAPHIC, col_lat, col_lon, lat, lon, yo,
                       var, float(np.float32(1.05)), radius)
    _, _, nloc_o = orc.letkf_standard(prob, X, Yb)
    assert nloc_o.sum() >= 6 * 3 * 2  # the inside points of every boundary set
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, col_lat, col_lon)
    obs = mda.Observation(yo, var, lat, lon, coord=mda._lib.COORD_GEOGRAPHIC)
    cfg = mda.Config({"inflation": 1.05, "localization_radius": radius})
    for prec in (mda._lib.PREC_FP64, mda._lib.PREC_BF16X3):
        ens = mda.Ensemble(X.copy(), nlon, nlat, nlev)
        f = mda.LETKF(ens, obs, mda.ObsOperator(hx=Yb), cfg, geometry=geom, ctx=ctx,
                      precision=prec)
        f.Analyse()
        np.testing.assert_array_equal(f.local_counts(), nloc_o)
    ens = mda.Ensemble(X.copy(), nlon, nlat, nlev)
    f = mda.LETKF(ens, obs, mda.ObsOperator(hx=Yb), cfg, geometry=geom, ctx=ctx,
                  form="refcompat")
    f.Analyse()
    np.testing.assert_array_equal(f.local_counts(), nloc_o)
    Xo = orc.letkf_refcompat_snapshot(prob, X, Yb)
    assert rel_err(ens.data, Xo) < TOL


# ---- invalid observations: variance +inf (GridObservation.hpp:239-252) ---------------------
def test_letkf_invalid_observations_carry_no_weight(mda, ctx, orc):
    case = latlon_case(nlon=32, nlat=16, nlev=3, k=12, p=900, seed=31)
    var = case["var"].copy()
    bad = np.zeros(case["p"], bool)
    bad[::7] = True
    var[bad] = np.inf
    Yb = case["H"].apply_host(case["X"])
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    cfg = mda.Config({"inflation": 1.05, "localization_radius": 2500.0})

    def run(sel):
        ens = mda.Ensemble(case["X"].copy(), 32, 16, 3)
        obs = mda.Observation(case["yo"][sel], var[sel], case["lat"][sel], case["lon"][sel],
                              coord=mda._lib.COORD_GEOGRAPHIC)
        f = mda.LETKF(ens, obs, mda.ObsOperator(hx=np.ascontiguousarray(Yb[:, sel])), cfg,
                      geometry=geom, ctx=ctx)
        f.Analyse()
        return ens.data
    all_obs = run(np.ones(case["p"], bool))
    valid_only = run(~bad)
    assert rel_err(all_obs, valid_only) < 1e-12
    prob = orc.Problem(12, 32, 16, 3, orc.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"],
                       case["lat"], case["lon"], case["yo"], var, np.float32(1.05).item(),
                       2500.0)
    Xo, _, _ = orc.letkf_standard(prob, case["X"], Yb)
    assert rel_err(all_obs, Xo) < TOL


def test_etkf_enkf_invalid_observations_carry_no_weight(mda, ctx, orc):
    case = l96_case(orc, N=40, k=20, seed=9)
    var = case["var"].copy()
    bad = np.zeros(case["p"], bool)
    bad[[3, 11]] = True
    var[bad] = np.inf
    H = mda.ObsOperator(case["rp"], case["cols"], case["ws"])
    Hv = mda.ObsOperator(*_csr_rows(case, ~bad))

    def run(cls, sel, Hop, **kw):
        ens = mda.Ensemble(case["X"].copy(), case["N"], 1, 1)
        obs = mda.Observation(case["yo"][sel], var[sel], case["obs_x"][sel], case["obs_y"][sel])
        if cls == "etkf":
            mda.ETKF(ens, obs, Hop, mda.Config({"inflation": 1.05}), ctx=ctx).Analyse()
        else:
            f = mda.EnKF(ens, obs, Hop, mda.Config({"analysis": {"inflation": 1.05}}), ctx=ctx)
            f.Analyse(Z=kw["Z"][sel])
        return ens.data
    full = np.ones(case["p"], bool)
    assert rel_err(run("etkf", full, H), run("etkf", ~bad, Hv)) < 1e-12
    assert rel_err(run("enkf", full, H, Z=case["Z"]), run("enkf", ~bad, Hv, Z=case["Z"])) < 1e-12


def _csr_rows(case, sel):
    rp, cols, ws = case["rp"], case["cols"], case["ws"]
    nrp, ncols, nws = [0], [], []
    for o in np.flatnonzero(sel):
        ncols += list(cols[rp[o]:rp[o + 1]]); nws += list(ws[rp[o]:rp[o + 1]])
        nrp.append(len(ncols))
    return np.array(nrp), np.array(ncols), np.array(nws)


# ---- non-finite inputs, k < 2, p = 0 ------------------------------------------------------------
def _dev_letkf(mda, ctx, case, X, yo, form="standard"):
    import torch
    Xd = torch.from_numpy(X.copy()).to("cuda:0")
    ens = mda.Ensemble(Xd, case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(yo, case["var"], case["lat"], case["lon"],
                          coord=mda._lib.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                   "localization_radius": 2500.0}),
                  geometry=geom, ctx=ctx, form=form)
    return f, Xd


def test_nonfinite_inputs_are_reported(mda, ctx):
    from metada_amd._lib import MdaError, MDA_E_NUMERIC
    case = latlon_case(nlon=16, nlat=8, nlev=2, k=6, p=200, seed=2)
    X = case["X"]
    # a NaN observation value: MDA_E_NUMERIC, device ensemble unchanged
    yo = case["yo"].copy()
    yo[17] = np.nan
    f, Xd = _dev_letkf(mda, ctx, case, X, yo)
    with pytest.raises(MdaError) as e:
        f.Analyse()
    assert e.value.code == MDA_E_NUMERIC
    np.testing.assert_array_equal(Xd.cpu().numpy(), X)
    # a NaN background value inside an observation's stencil: H(x) is
    # non-finite, so the ensemble is left unchanged as well
    tap = int(case["H"].col[case["H"].row_ptr[5]])
    Xn = X.copy()
    Xn[3, tap] = np.nan
    f, Xd = _dev_letkf(mda, ctx, case, Xn, case["yo"])
    with pytest.raises(MdaError) as e:
        f.Analyse()
    assert e.value.code == MDA_E_NUMERIC
    np.testing.assert_array_equal(np.isnan(Xd.cpu().numpy()), np.isnan(Xn))
    np.testing.assert_array_equal(np.nan_to_num(Xd.cpu().numpy()), np.nan_to_num(Xn))
    # an infinite background value at a point no observation reads: reported,
    # and the analysis there is non-finite (as the reference's arithmetic
    # makes it)
    used = set(int(c) for c in case["H"].col)
    free = next(i for i in range(X.shape[1]) if i not in used)
    Xi = X.copy()
    Xi[0, free] = np.inf
    f, Xd = _dev_letkf(mda, ctx, case, Xi, case["yo"])
    with pytest.raises(MdaError, match="non-finite background"):
        f.Analyse()
    assert not np.isfinite(Xd.cpu().numpy()[:, free]).all()
    # a variance that is not > 0
    obs = mda.Observation(case["yo"], np.where(np.arange(case["p"]) == 4, 0.0, case["var"]),
                          case["lat"], case["lon"], coord=mda._lib.COORD_GEOGRAPHIC)
    ens = mda.Ensemble(X.copy(), 16, 8, 2)
    with pytest.raises(MdaError):
        mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                   "localization_radius": 2500.0}),
                  geometry=mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"],
                                              case["col_lon"]), ctx=ctx).Analyse()
    np.testing.assert_array_equal(ens.data, X)


def test_k_below_two_and_empty_observations(mda, ctx, orc):
    from metada_amd._lib import MdaError, MDA_E_ARG
    rng = np.random.default_rng(3)
    X1 = rng.standard_normal((1, 30))
    obs = mda.Observation(np.zeros(2), np.ones(2), np.array([1.0, 4.0]), np.zeros(2))
    H = mda.ObsOperator(np.array([0, 1, 2]), np.array([1, 4]), np.ones(2))
    for algo in ("letkf", "etkf"):
        ens = mda.Ensemble(X1.copy(), 30, 1, 1)
        with pytest.raises(MdaError) as e:
            if algo == "letkf":
                mda.LETKF(ens, obs, H, mda.Config({"inflation": 1.0, "localization_radius": 2.0}),
                          ctx=ctx).Analyse()
            else:
                mda.ETKF(ens, obs, H, mda.Config({"inflation": 1.0}), ctx=ctx).Analyse()
        assert e.value.code == MDA_E_ARG
    # p = 0: the LETKF gives every point the empty-column update
    # x_a = xbar + sqrt(rho)(x - xbar) (LETKF.hpp:167-190) ...
    nx, ny, nlev, k = 6, 4, 2, 5
    X = rng.standard_normal((k, nx * ny * nlev))
    empty = mda.Observation(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
    H0 = mda.ObsOperator(np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0))
    ens = mda.Ensemble(X.copy(), nx, ny, nlev)
    f = mda.LETKF(ens, empty, H0, mda.Config({"inflation": 1.21, "localization_radius": 2.0}),
                  ctx=ctx)
    d = f.Analyse()
    assert d["n_empty_columns"] == nx * ny and d["total_local_obs"] == 0
    rho = float(np.float32(1.21))
    want = X.mean(0) + math.sqrt(rho) * (X - X.mean(0))
    assert rel_err(ens.data, want) < 1e-14
    prob = orc.Problem(k, nx, ny, nlev, orc.COORD_GRID, np.tile(np.arange(nx, dtype=float), ny),
                       np.repeat(np.arange(ny, dtype=float), nx), np.zeros(0), np.zeros(0),
                       np.zeros(0), np.zeros(0), rho, 2.0)
    Xo, _, _ = orc.letkf_standard(prob, X, np.zeros((k, 0)))
    assert rel_err(ens.data, Xo) < TOL
    # ... while the global filters need observations (their S / Pa would be
    # empty; the reference's Eigen calls are undefined there)
    ens = mda.Ensemble(X.copy(), nx * ny * nlev, 1, 1)
    with pytest.raises(MdaError) as e:
        mda.ETKF(ens, empty, H0, mda.Config({"inflation": 1.0}), ctx=ctx).Analyse()
    assert e.value.code == MDA_E_ARG


def test_non_spd_leaves_device_ensemble_unchanged(mda, ctx):
    """Overflowing H(x) makes Y'^T Y' infinite, so the Cholesky factor of
    the reference form's Pa (refcompat LETKF, ETKF) does not exist: the call
    returns MDA_E_NUMERIC and the device-resident ensemble is untouched."""
    import torch
    from metada_amd._lib import MdaError, MDA_E_NUMERIC
    rng = np.random.default_rng(8)
    nx, k, p = 40, 8, 6
    X = rng.standard_normal((k, nx))
    Yb = rng.standard_normal((k, p)) * 1e160
    obs = mda.Observation(np.zeros(p), np.ones(p), np.arange(p, dtype=float) * 3, np.zeros(p))
    for algo in ("letkf", "etkf"):
        Xd = torch.from_numpy(X.copy()).to("cuda:0")
        ens = mda.Ensemble(Xd, nx, 1, 1)
        with pytest.raises(MdaError) as e:
            if algo == "letkf":
                mda.LETKF(ens, obs, mda.ObsOperator(hx=Yb),
                          mda.Config({"inflation": 1.0, "localization_radius": 50.0}), ctx=ctx,
                          form="refcompat").Analyse()
            else:
                mda.ETKF(ens, obs, mda.ObsOperator(hx=Yb), mda.Config({"inflation": 1.0}),
                         ctx=ctx).Analyse()
        assert e.value.code == MDA_E_NUMERIC
        np.testing.assert_array_equal(Xd.cpu().numpy(), X)


# ---- EnKF beyond k = 64 and p = 128 (EnKF.hpp:188-209) ------------------------------------------
@pytest.mark.parametrize("k,p,hetero", [(96, 60, True), (20, 300, True), (24, 200, False)])
def test_enkf_large_k_and_condition_number(mda, ctx, orc, k, p, hetero):
    rng = np.random.default_rng(k + p)
    n = 150
    X = rng.standard_normal((k, n)) + np.linspace(-1, 1, n)
    idx = rng.integers(0, n, p)
    rp = np.arange(p + 1)
    H = mda.ObsOperator(rp, idx.astype(np.int64), np.ones(p))
    var = rng.uniform(0.2, 2.0, p) if hetero else np.full(p, 0.64)
    yo = X.mean(0)[idx] + rng.standard_normal(p)
    Z = rng.standard_normal((p, k))
    ens = mda.Ensemble(X.copy(), n, 1, 1)
    obs = mda.Observation(yo, var, idx.astype(float), np.zeros(p))
    f = mda.EnKF(ens, obs, H, mda.Config({"analysis": {"inflation": 1.1}}), ctx=ctx)
    f.Analyse(Z=Z)
    g = f.getAnalysisResults()
    Yb = X[:, idx]
    Xo, d = orc.enkf_analyse(X, Yb, yo, var, Z, float(np.float32(1.1)))
    assert rel_err(ens.data, Xo) < TOL
    for key in ("innovation_norm", "background_spread", "analysis_spread", "max_kalman_gain",
                "min_kalman_gain", "condition_number"):
        assert abs(g[key] - d[key]) <= 1e-10 * abs(d[key]), key
