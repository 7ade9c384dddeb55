"""The eigenvector kernel with the transform applied in place (k in {16, 32,
48, 64}, keep_transforms = 0) against the path that stores every column's T
and transforms afterwards (keep_transforms = 1) and against the oracle
(LETKF.hpp:213-242 in Hunt's form): bit-identical ensembles, with partial
groups of 4 columns, several 16-level chunks, column passes, per-column rho,
empty columns and a column left to the Jacobi fallback."""
import numpy as np
import pytest

from cases import latlon_case, rel_err

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


def _run_latlon(mda, ctx, case, radius, rho, keep, precision=None, device=False):
    X = case["X"].copy()
    if device:
        import torch
        X = torch.from_numpy(X).cuda()
    ens = mda.Ensemble(X, case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                          coord=mda._lib.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    kw = {} if precision is None else {"precision": precision}
    f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": rho,
                                                   "localization_radius": radius}),
                  geometry=geom, ctx=ctx, keep_transforms=keep, **kw)
    diag = f.Analyse()
    return ens.to_host(), f, diag


DIAG_KEYS = ("n_columns", "n_empty_columns", "total_local_obs", "max_local_obs",
             "n_fallback_columns", "max_jacobi_sweeps")


@pytest.mark.parametrize("k", [16, 32, 48, 64])
def test_fused_equals_kept_and_oracle(mda, ctx, orc, k):
    """63 x 31 columns (the last group of 4 holds one), 20 levels (a full and a
    partial chunk), clustered observations (empty columns)."""
    case = latlon_case(nlon=63, nlat=31, nlev=20, k=k, p=2500, seed=70 + k, cluster=True,
                       nmodes=40)
    Xf, ff, df = _run_latlon(mda, ctx, case, 1500.0, 1.07, keep=False)
    cols = np.array([0, 31, 1952])
    # the fused path stores no T
    with pytest.raises(mda._lib.MdaError) as e:
        ff.transforms(cols)
    assert "keep_transforms" in str(e.value)
    Xk, fk, dk = _run_latlon(mda, ctx, case, 1500.0, 1.07, keep=True)
    np.testing.assert_array_equal(Xf, Xk)
    for key in DIAG_KEYS:
        assert df[key] == dk[key], key
    assert df["n_empty_columns"] > 0
    prob = orc.Problem(k, case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], ff.inflation, ff.localization_radius)
    Xo, To, _ = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]), want_T=True)
    assert rel_err(Xf, Xo) < TOL
    assert rel_err(fk.transforms(cols), To[cols]) < TOL


def test_fused_device_tensor_and_bf16x3_gram(mda, ctx, orc):
    case = latlon_case(nlon=40, nlat=20, nlev=6, k=64, p=3000, seed=5)
    Xf, _, _ = _run_latlon(mda, ctx, case, 2000.0, 1.05, keep=False, device=True)
    Xk, _, _ = _run_latlon(mda, ctx, case, 2000.0, 1.05, keep=True)
    np.testing.assert_array_equal(Xf, Xk)
    code = mda._lib.PREC_BF16X3
    Xb, _, _ = _run_latlon(mda, ctx, case, 2000.0, 1.05, keep=False, precision=code)
    Xbk, _, _ = _run_latlon(mda, ctx, case, 2000.0, 1.05, keep=True, precision=code)
    np.testing.assert_array_equal(Xb, Xbk)


def test_fused_fallback_and_empty_columns(mda, ctx, orc):
    """Six columns of one row, k = 16, radius 0.5 grid units (each observation
    reaches its own column only).  Column 0 sees two co-located observations
    whose anomaly rows are orthogonal with equal norms (a double eigenvalue
    of A away from (k-1)/rho: the orthogonality check sends it to Jacobi);
    column 1 sees none (T = sqrt(rho) I); columns 2-5 see random ones.  The
    fused kernel updates columns 1-5, the restricted transform column 0."""
    k, nx, nlev = 16, 6, 3
    rng = np.random.default_rng(8)
    X = rng.standard_normal((k, nx * nlev))
    ox = [0.0, 0.0] + [float(c) for c in (2, 3, 3, 4, 5, 5, 5)]
    p = len(ox)
    hx = rng.standard_normal((k, p))
    hx[:, 0] = 0.0
    hx[:, 1] = 0.0
    hx[0:2, 0] = [1.0, -1.0]
    hx[2:4, 1] = [1.0, -1.0]
    yo = rng.standard_normal(p) * 0.3
    var = np.full(p, 0.25)
    out = {}
    for keep in (False, True):
        obs = mda.Observation(yo, var, np.array(ox), np.zeros(p), np.zeros(p))
        ens = mda.Ensemble(X.copy(), nx, 1, nlev)
        f = mda.LETKF(ens, obs, mda.ObsOperator(hx=hx),
                      mda.Config({"inflation": 1.05, "localization_radius": 0.5}), ctx=ctx,
                      keep_transforms=keep)
        d = f.Analyse()
        out[keep] = (ens.data.copy(), d, f)
    np.testing.assert_array_equal(out[False][0], out[True][0])
    d = out[False][1]
    assert d["n_fallback_columns"] == 1 and d["n_empty_columns"] == 1
    assert out[False][2].column_flags()[0] != 0
    prob = orc.Problem(k, nx, 1, nlev, orc.COORD_GRID, np.arange(nx, dtype=float),
                       np.zeros(nx), np.array(ox), np.zeros(p), yo, var,
                       out[False][2].inflation, out[False][2].localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, X, hx)
    assert rel_err(out[False][0], Xo) < TOL


def test_fused_per_column_rho_matches_oracle(mda, ctx, orc):
    rng = np.random.default_rng(41)
    nx, ny, nlev, k, p = 21, 13, 17, 32, 260
    X = rng.standard_normal((k, nx * ny * nlev))
    ox, oy = rng.uniform(0, nx - 1, p), rng.uniform(0, ny - 1, p)
    oz = rng.integers(0, nlev, p).astype(float)
    obs = mda.Observation(rng.standard_normal(p), rng.uniform(0.2, 2.0, p), ox, oy, oz)
    H = mda.ObsOperator.identity_idw4(obs, nx, ny, nlev)
    rho_col = rng.uniform(1.0, 1.3, nx * ny)
    res = []
    for keep in (False, True):
        ens = mda.Ensemble(X.copy(), nx, ny, nlev)
        f = mda.LETKF(ens, obs, H, mda.Config({"inflation": 1.05, "localization_radius": 4.5}),
                      ctx=ctx, rho_column=rho_col, keep_transforms=keep)
        f.Analyse()
        res.append(ens.data.copy())
    np.testing.assert_array_equal(res[0], res[1])
    prob = orc.Problem(k, nx, ny, nlev, orc.COORD_GRID, np.tile(np.arange(nx, dtype=float), ny),
                       np.repeat(np.arange(ny, dtype=float), nx), ox, oy, obs.values,
                       obs.variances, f.inflation, f.localization_radius, rho_col=rho_col)
    Xo, _, _ = orc.letkf_standard(prob, X, H.apply_host(X))
    assert rel_err(res[0], Xo) < TOL


@pytest.mark.parametrize("nlev", [1, 16, 32, 48, 53, 64])
def test_fused_level_chunks(mda, ctx, orc, nlev):
    """Level counts at and around the 16-level chunks and the two chunks the
    eigenvector kernel loads ahead (32): the fused path against the
    stored-T path (bit for bit) and the oracle, on 9 x 7 columns (a partial
    group of four)."""
    case = latlon_case(nlon=9, nlat=7, nlev=nlev, k=32, p=400, seed=300 + nlev, nmodes=20)
    Xf, ff, _ = _run_latlon(mda, ctx, case, 3000.0, 1.05, keep=False)
    Xk, _, _ = _run_latlon(mda, ctx, case, 3000.0, 1.05, keep=True)
    np.testing.assert_array_equal(Xf, Xk)
    prob = orc.Problem(32, case["nlon"], case["nlat"], nlev, orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], ff.inflation, ff.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    assert rel_err(Xf, Xo) < TOL


@pytest.mark.parametrize("nlev", [3, 20])
def test_k128_transform_odd_columns(mda, ctx, orc, nlev, monkeypatch):
    """k = 128 on 15 x 7 columns (the last two-column block holds one) with
    one and two 16-level chunks: the two-column MFMA transform against the
    oracle, in one pass and in passes of 40 columns (pass offsets, a short
    last pass), bit for bit between the two."""
    case = latlon_case(nlon=15, nlat=7, nlev=nlev, k=128, p=2500, seed=77 + nlev, nmodes=60)
    monkeypatch.delenv("MDA_LETKF_CHUNK_COLS", raising=False)
    Xg, f, d = _run_latlon(mda, ctx, case, 4000.0, 1.05, keep=False)
    monkeypatch.setenv("MDA_LETKF_CHUNK_COLS", "40")
    Xp, _, _ = _run_latlon(mda, ctx, case, 4000.0, 1.05, keep=False)
    monkeypatch.delenv("MDA_LETKF_CHUNK_COLS", raising=False)
    np.testing.assert_array_equal(Xg, Xp)
    prob = orc.Problem(128, case["nlon"], case["nlat"], nlev, orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], f.inflation, f.localization_radius)
    Xo, _, nloc = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    assert d["total_local_obs"] == nloc.sum()
    assert rel_err(Xg, Xo) < TOL
