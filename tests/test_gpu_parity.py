"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded inputs.  FP64 gate: 1e-10 relative (max-norm)."""
import os

import numpy as np
import pytest

from cases import l96_case, latlon_case, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
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


def test_mean_perturbations_match_reference_order(mda, ctx, orc):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((13, 5001)) * 3 + 1
    ens = mda.Ensemble(X, 5001, 1, 1)
    m = ens.RecomputeMean(ctx)
    np.testing.assert_array_equal(m, orc.ensemble_mean(X))       # same summation order
    P = ens.Perturbations(np.sqrt(1.05), ctx)
    np.testing.assert_allclose(P, orc.ensemble_perturbations(X, m, np.sqrt(1.05)), rtol=1e-15,
                               atol=1e-15)


def test_obs_innovations_csr_and_hx(mda, ctx, orc):
    import ctypes as C
    case = latlon_case(nlon=32, nlat=16, nlev=3, k=11, p=777)
    X, H = case["X"], case["H"]
    ens = mda.Ensemble(X, 32, 16, 3)
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"])
    from metada_amd.analysis import _obs_struct
    es, ek = ens.struct(); os_, ok = _obs_struct(obs); hs, hk = H.struct()
    k, p = 11, 777
    ybar, yp, d = np.zeros(p), np.zeros((k, p)), np.zeros(p)
    ctx.check(ctx.lib.mda_obs_innovations(ctx.h, C.byref(es), C.byref(os_), C.byref(hs),
                                          ybar.ctypes.data, yp.ctypes.data, d.ctypes.data, 0))
    Yb = np.stack([orc.csr_apply(H.row_ptr, H.col, H.weight, X[j]) for j in range(k)])
    assert rel_err(ybar, Yb.mean(0)) < 1e-14
    assert rel_err(yp, Yb - Yb.mean(0)) < 1e-13
    assert rel_err(d, case["yo"] - Yb.mean(0)) < 1e-13
    # caller-supplied H(x)
    H2 = mda.ObsOperator(hx=Yb)
    hs2, hk2 = H2.struct()
    yp2 = np.zeros((k, p))
    ctx.check(ctx.lib.mda_obs_innovations(ctx.h, C.byref(es), C.byref(os_), C.byref(hs2),
                                          None, yp2.ctypes.data, None, 0))
    assert rel_err(yp2, Yb - Yb.mean(0)) < 1e-14


def _enkf_setup(mda, case):
    ens = mda.Ensemble(case["X"].copy(), case["N"], 1, 1)
    obs = mda.Observation(case["yo"], case["var"], case["obs_x"], case["obs_y"])
    H = mda.ObsOperator(case["rp"], case["cols"], case["ws"])
    return ens, obs, H


def test_enkf_l96_matches_oracle(mda, ctx, orc):
    case = l96_case(orc)
    ens, obs, H = _enkf_setup(mda, case)
    cfg = mda.Config({"analysis": {"inflation": 1.05, "inflation_method": "multiplicative",
                                   "output_base_file": "x", "format": "txt"}})
    f = mda.EnKF(ens, obs, H, cfg, ctx=ctx)
    f.Analyse(Z=case["Z"])
    Yb = np.stack([orc.csr_apply(case["rp"], case["cols"], case["ws"], x) for x in case["X"]])
    Xo, dg = orc.enkf_analyse(case["X"], Yb, case["yo"], case["var"], case["Z"],
                              f.inflation_factor)
    assert f.inflation_factor == 1.0499999523162842  # float-narrowed as the reference
    assert rel_err(ens.data, Xo) < TOL
    res = f.getAnalysisResults()
    for key in ("innovation_norm", "background_spread", "analysis_spread",
                "max_kalman_gain", "min_kalman_gain", "condition_number"):
        assert abs(res[key] - dg[key]) <= TOL * abs(dg[key]), key
    assert res["analysis_increment_norm"] == 0.0
    assert res["ensemble_size"] == 20 and res["observation_count"] == 20


def test_enkf_large_p_constant_R_condition(mda, ctx, orc):
    # p > 128 with R = sigma^2 I: cond(S) from the k x k eigenproblem
    rng = np.random.default_rng(3)
    k, n, p = 10, 400, 300
    X = rng.standard_normal((k, n))
    cols = rng.integers(0, n, p)
    rp = np.arange(p + 1); ws = np.ones(p)
    yo = rng.standard_normal(p); var = np.full(p, 0.7); Z = rng.standard_normal((p, k))
    ens = mda.Ensemble(X.copy(), n, 1, 1)
    obs = mda.Observation(yo, var, np.zeros(p), np.zeros(p))
    f = mda.EnKF(ens, obs, mda.ObsOperator(rp, cols, ws),
                 mda.Config({"analysis": {"inflation": 1.0}}), ctx=ctx)
    f.Analyse(Z=Z)
    Yb = X[:, cols]
    Xo, dg = orc.enkf_analyse(X, Yb, yo, var, Z, 1.0)
    assert rel_err(ens.data, Xo) < TOL
    assert abs(f.getAnalysisResults()["condition_number"] - dg["condition_number"]) < 1e-9 * dg["condition_number"]


def test_etkf_matches_oracle(mda, ctx, orc):
    case = l96_case(orc, k=16, seed=9)
    ens, obs, H = _enkf_setup(mda, case)
    f = mda.ETKF(ens, obs, H, mda.Config({"inflation": 1.1}), ctx=ctx)
    f.Analyse()
    Yb = np.stack([orc.csr_apply(case["rp"], case["cols"], case["ws"], x) for x in case["X"]])
    Xo = orc.etkf_analyse(case["X"], Yb, case["yo"], case["var"], f.inflation)
    assert rel_err(ens.data, Xo) < TOL


def _letkf_run(mda, ctx, case, radius, rho, form="standard", device=False, cols_T=None):
    nlon, nlat, nlev = case["nlon"], case["nlat"], case["nlev"]
    X = case["X"].copy()
    if device:
        import torch
        X = torch.from_numpy(X).cuda()
    ens = mda.Ensemble(X, nlon, nlat, nlev)
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                          coord=mda._lib.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    cfg = mda.Config({"inflation": rho, "localization_radius": radius})
    f = mda.LETKF(ens, obs, case["H"], cfg, geometry=geom, ctx=ctx, form=form,
                  keep_transforms=cols_T is not None)
    diag = f.Analyse()
    T = f.transforms(cols_T) if cols_T is not None else None
    return ens.to_host(), f, diag, T


def _oracle_problem(orc, case, f):
    return orc.Problem(case["k"], case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], f.inflation, f.localization_radius)


@pytest.mark.parametrize("k", [20, 64, 7])
def test_letkf_standard_matches_oracle(mda, ctx, orc, k):
    case = latlon_case(k=k, p=3000 if k != 7 else 800, seed=11 + k)
    cols = np.array([0, 5, 100, 1000, 2047])
    Xg, f, diag, Tg = _letkf_run(mda, ctx, case, 2000.0, 1.05, cols_T=cols)
    prob = _oracle_problem(orc, case, f)
    Yb = case["H"].apply_host(case["X"])
    Xo, To, nloc = orc.letkf_standard(prob, case["X"], Yb, want_T=True)
    assert rel_err(Xg, Xo) < TOL
    assert rel_err(Tg, To[cols]) < TOL
    assert diag["total_local_obs"] == nloc.sum()
    assert diag["max_local_obs"] == nloc.max()
    assert diag["n_empty_columns"] == int((nloc == 0).sum())


def test_letkf_standard_device_tensor_in_place(mda, ctx, orc):
    case = latlon_case(k=20, p=2500, seed=5, cluster=True)
    Xg, f, diag, _ = _letkf_run(mda, ctx, case, 1500.0, 1.1, device=True)
    prob = _oracle_problem(orc, case, f)
    Xo, _, nloc = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    assert diag["n_empty_columns"] > 0  # clustered obs leave columns without local obs
    assert rel_err(Xg, Xo) < TOL


def test_letkf_refcompat_tutorial_matches_oracle(mda, ctx, orc):
    """The reference's tutorial LETKF case (36 x 18 grid, 9 members, 28 grid
    obs, radius 10, inflation 1.0; src/backends/simple/tutorial/letkf.yaml)
    in the reference semantics with snapshot H."""
    tut = os.path.join(ROOT, "tests", "golden", "tutorial")
    X = np.stack([mda.read_simple_state(os.path.join(tut, f"ens_{i}.txt")) for i in range(1, 10)])
    obs = mda.Observation.from_grid_file(os.path.join(tut, "obs_format.txt"), error=np.float32(0.1))
    H = mda.ObsOperator.identity_idw4(obs, 36, 18, 1)
    cfg = mda.Config({"inflation": 1.0, "localization_radius": 10.0})
    ens = mda.Ensemble(X.copy(), 36, 18, 1)
    f = mda.LETKF(ens, obs, H, cfg, ctx=ctx, form="refcompat")
    f.Analyse()
    prob = orc.Problem(9, 36, 18, 1, orc.COORD_GRID,
                       np.tile(np.arange(36.0), 18), np.repeat(np.arange(18.0), 36),
                       obs.loc_a, obs.loc_b, obs.values, obs.variances, 1.0, 10.0)
    Xo = orc.letkf_refcompat_snapshot(prob, X, H.apply_host(X))
    assert rel_err(ens.data, Xo) < TOL


def test_letkf_refcompat_latlon_with_inflation(mda, ctx, orc):
    case = latlon_case(k=16, p=1500, seed=21)
    Xg, f, diag, _ = _letkf_run(mda, ctx, case, 1200.0, 1.2, form="refcompat")
    prob = _oracle_problem(orc, case, f)
    Xo = orc.letkf_refcompat_snapshot(prob, case["X"], case["H"].apply_host(case["X"]))
    assert rel_err(Xg, Xo) < TOL


def test_letkf_grid_coordinates_and_per_column_rho(mda, ctx, orc):
    rng = np.random.default_rng(4)
    nx, ny, nlev, k, p = 24, 17, 2, 12, 150
    X = rng.standard_normal((k, nx * ny * nlev))
    ox, oy = rng.uniform(0, nx - 1, p), rng.uniform(0, ny - 1, p)
    oz = rng.integers(0, nlev, p).astype(float)
    obs = mda.Observation(rng.standard_normal(p), rng.uniform(0.2, 2.0, p), ox, oy, oz)
    H = mda.ObsOperator.identity_idw4(obs, nx, ny, nlev)
    rho_col = rng.uniform(1.0, 1.3, nx * ny)
    ens = mda.Ensemble(X.copy(), nx, ny, nlev)
    f = mda.LETKF(ens, obs, H, mda.Config({"inflation": 1.05, "localization_radius": 4.5}),
                  ctx=ctx, rho_column=rho_col)
    f.Analyse()
    prob = orc.Problem(k, nx, ny, nlev, orc.COORD_GRID, np.tile(np.arange(nx, dtype=float), ny),
                       np.repeat(np.arange(ny, dtype=float), nx), ox, oy, obs.values,
                       obs.variances, f.inflation, f.localization_radius, rho_col=rho_col)
    Xo, _, _ = orc.letkf_standard(prob, X, H.apply_host(X))
    assert rel_err(ens.data, Xo) < TOL


def test_metrics_match_oracle(mda, ctx, orc):
    rng = np.random.default_rng(8)
    X = rng.standard_normal((9, 3000)); t = rng.standard_normal(3000)
    g = mda.metrics(mda.Ensemble(X, 3000, 1, 1), t, ctx)
    o = orc.metrics(X, t)
    for key in ("rmse", "bias", "correlation", "crps", "avg_spread"):
        assert abs(g[key] - o[key]) <= 1e-12 * max(1.0, abs(o[key])), key


def test_error_paths(mda, ctx):
    from metada_amd._lib import MdaError
    rng = np.random.default_rng(0)
    X = rng.standard_normal((4, 10))
    obs = mda.Observation(np.zeros(2), np.ones(2), np.zeros(2), np.zeros(2))
    bad_H = mda.ObsOperator(np.array([0, 1, 2]), np.array([0, 99]), np.ones(2))
    with pytest.raises(MdaError, match="MDA_E_ARG"):
        mda.ETKF(mda.Ensemble(X.copy(), 10, 1, 1), obs, bad_H, mda.Config({"inflation": 1.0}),
                 ctx=ctx).Analyse()
    H = mda.ObsOperator(np.array([0, 1, 2]), np.array([0, 3]), np.ones(2))
    with pytest.raises(MdaError, match="MDA_E_ARG"):
        mda.LETKF(mda.Ensemble(X.copy(), 10, 1, 1), obs, H,
                  mda.Config({"inflation": 1.0, "localization_radius": -1.0}), ctx=ctx).Analyse()


@pytest.mark.parametrize("k", [64, 20])
def test_letkf_tridiagonal_eigensolver_matches_jacobi(mda, ctx, orc, k):
    """The default symmetric square root (tridiagonal + twisted factorisation)
    against the parallel Jacobi path and the oracle on the same inputs.  The
    fields carry many modes so Y'^T R^-1 Y' has a well-separated spectrum, as
    Gaussian random fields do."""
    case = latlon_case(k=k, p=6000, seed=40 + k, nmodes=96)
    out = {}
    for name, meth in (("auto", mda._lib.EIG_AUTO), ("jacobi", mda._lib.EIG_JACOBI)):
        ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
        obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                              coord=mda._lib.COORD_GEOGRAPHIC)
        geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
        f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                       "localization_radius": 1800.0}),
                      geometry=geom, ctx=ctx, eig_method=meth)
        diag = f.Analyse()
        out[name] = (ens.data.copy(), diag)
    # the odd column with a close, significant eigenpair goes to Jacobi
    assert out["auto"][1]["n_fallback_columns"] <= 0.01 * case["nlon"] * case["nlat"]
    assert out["jacobi"][1]["max_jacobi_sweeps"] > 0
    assert rel_err(out["auto"][0], out["jacobi"][0]) < 1e-11
    prob = orc.Problem(k, case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], f.inflation, f.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    assert rel_err(out["auto"][0], Xo) < TOL


def test_letkf_rank_deficient_gram_needs_no_fallback(mda, ctx, orc):
    """Few local observations make Y'^T R^-1 Y' rank deficient, so A has the
    eigenvalue (k-1)/rho with high multiplicity; W is formed around it, so the
    tridiagonal path handles these columns itself."""
    rng = np.random.default_rng(12)
    nx, ny, k, p = 10, 8, 16, 12
    X = rng.standard_normal((k, nx * ny))
    ox, oy = rng.uniform(0, nx - 1, p), rng.uniform(0, ny - 1, p)
    obs = mda.Observation(rng.standard_normal(p), np.full(p, 0.5), ox, oy, np.zeros(p))
    H = mda.ObsOperator.identity_idw4(obs, nx, ny, 1)
    ens = mda.Ensemble(X.copy(), nx, ny, 1)
    f = mda.LETKF(ens, obs, H, mda.Config({"inflation": 1.1, "localization_radius": 2.5}),
                  ctx=ctx)
    diag = f.Analyse()
    assert diag["n_fallback_columns"] == 0
    prob = orc.Problem(k, nx, ny, 1, orc.COORD_GRID, np.tile(np.arange(nx, dtype=float), ny),
                       np.repeat(np.arange(ny, dtype=float), nx), ox, oy, obs.values,
                       obs.variances, f.inflation, f.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, X, H.apply_host(X))
    assert rel_err(ens.data, Xo) < TOL


@pytest.mark.parametrize("k", [8, 128])
def test_letkf_repeated_eigenvalue_falls_back_to_jacobi(mda, ctx, orc, k):
    """Two co-located observations whose anomaly rows are orthogonal with equal
    norms give A a double eigenvalue away from (k-1)/rho: the twisted-
    factorisation vectors coincide, the orthogonality check flags the column
    and the Jacobi kernel re-solves it (k = 128: the four-wavefront
    eigenvector kernel and the Jacobi kernel in global scratch)."""
    rng = np.random.default_rng(2)
    X = rng.standard_normal((k, 3))
    hx = np.zeros((k, 2))
    hx[0:2, 0] = [1, -1]
    hx[2:4, 1] = [1, -1]
    yo = np.array([0.3, -0.2])
    obs = mda.Observation(yo, np.full(2, 0.25), np.zeros(2), np.zeros(2), np.zeros(2))
    ens = mda.Ensemble(X.copy(), 1, 1, 3)
    f = mda.LETKF(ens, obs, mda.ObsOperator(hx=hx),
                  mda.Config({"inflation": 1.05, "localization_radius": 1.0}), ctx=ctx)
    diag = f.Analyse()
    assert diag["n_fallback_columns"] == 1
    prob = orc.Problem(k, 1, 1, 3, orc.COORD_GRID, [0.0], [0.0], np.zeros(2), np.zeros(2), yo,
                       np.full(2, 0.25), f.inflation, f.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, X, hx)
    assert rel_err(ens.data, Xo) < TOL


@pytest.mark.parametrize("prec,bound", [("fp32", 1e-6), ("bf16x3", 1e-4)])
@pytest.mark.parametrize("k", [64, 40])
def test_letkf_mixed_precision_contractions(mda, ctx, orc, prec, bound, k):
    """FP32 / BF16x3 Y'^T R^-1 Y' on MFMA with the FP64 eigensolve: the
    analysis ensemble stays within `bound` relative RMS of the oracle's FP64
    analysis, and the mixed and FP64 GPU paths see the same local counts."""
    case = latlon_case(k=k, p=6000, seed=70 + k, nmodes=96)
    code = {"fp32": mda._lib.PREC_FP32, "bf16x3": mda._lib.PREC_BF16X3}[prec]
    out = {}
    for name, pc in (("fp64", mda._lib.PREC_FP64), (prec, code)):
        ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
        obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                              coord=mda._lib.COORD_GEOGRAPHIC)
        geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
        f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                       "localization_radius": 1800.0}),
                      geometry=geom, ctx=ctx, precision=pc)
        diag = f.Analyse()
        out[name] = (ens.data.copy(), diag, f.local_counts())
    prob = orc.Problem(k, case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], f.inflation, f.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    assert rel_err(out["fp64"][0], Xo) < TOL
    np.testing.assert_array_equal(out["fp64"][2], out[prec][2])
    assert out[prec][1]["total_local_obs"] == out["fp64"][1]["total_local_obs"]
    Xm = out[prec][0]
    rms = np.sqrt(np.mean((Xm - Xo) ** 2)) / np.sqrt(np.mean(Xo ** 2))
    assert rms < bound, rms
    # the increment itself is reproduced, not just the background
    inc = np.sqrt(np.mean((Xo - case["X"]) ** 2))
    assert np.sqrt(np.mean((Xm - Xo) ** 2)) < 10 * bound * inc


@pytest.mark.parametrize("gram", ["fp64", "bf16x3"])
@pytest.mark.parametrize("k", [64, 48, 32])
def test_letkf_bf16x3_transform(mda, ctx, orc, gram, k):
    """X_a = xbar + X' T with split-precision bf16 products and FP32
    accumulation (transform_precision = BF16X3), alone or with the BF16x3
    Gram: within 1e-4 relative RMS of the oracle's FP64 analysis, the
    increment reproduced, and the mean update (which the FP64 xbar carries)
    no worse than the perturbations."""
    case = latlon_case(k=k, p=6000, seed=90 + k, nmodes=96)
    L = mda._lib
    ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                          coord=L.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(L.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                   "localization_radius": 1800.0}),
                  geometry=geom, ctx=ctx,
                  precision=L.PREC_BF16X3 if gram == "bf16x3" else L.PREC_FP64,
                  transform_precision=L.PREC_BF16X3)
    f.Analyse()
    prob = orc.Problem(k, case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], f.inflation, f.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    Xm = ens.data
    rms = np.sqrt(np.mean((Xm - Xo) ** 2)) / np.sqrt(np.mean(Xo ** 2))
    assert rms < 1e-4, rms
    inc = np.sqrt(np.mean((Xo - case["X"]) ** 2))
    assert np.sqrt(np.mean((Xm - Xo) ** 2)) < 1e-3 * inc
    em = np.sqrt(np.mean((Xm.mean(0) - Xo.mean(0)) ** 2))
    ep = np.sqrt(np.mean(((Xm - Xm.mean(0)) - (Xo - Xo.mean(0))) ** 2))
    assert em <= 2 * ep + 1e-14


def test_letkf_bf16x3_transform_refuses_unsupported_k(mda, ctx):
    case = latlon_case(k=40, p=500, seed=3)
    L = mda._lib
    ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                          coord=L.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(L.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                   "localization_radius": 1800.0}),
                  geometry=geom, ctx=ctx, transform_precision=L.PREC_BF16X3)
    with pytest.raises(L.MdaError) as e:
        f.Analyse()
    assert e.value.code == L.MDA_E_ARG
    np.testing.assert_array_equal(ens.data, case["X"])


@pytest.mark.parametrize("prec,bound", [("fp64", TOL), ("bf16x3", 1e-4)])
def test_letkf_supplied_hx_matches_stencil(mda, ctx, orc, prec, bound):
    """H(x) handed over member-major (read in input order, rows moved into
    bucket order afterwards) and the CSR stencil (statistics computed in
    bucket order) give the same local sets and the oracle's analysis."""
    k = 40
    case = latlon_case(k=k, p=5000, seed=91, nmodes=96)
    code = {"fp64": mda._lib.PREC_FP64, "bf16x3": mda._lib.PREC_BF16X3}[prec]
    hx = case["H"].apply_host(case["X"])
    out = {}
    for name, H in (("stencil", case["H"]), ("hx", mda.ObsOperator(hx=hx))):
        ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
        obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                              coord=mda._lib.COORD_GEOGRAPHIC)
        geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
        f = mda.LETKF(ens, obs, H, mda.Config({"inflation": 1.05,
                                               "localization_radius": 1800.0}),
                      geometry=geom, ctx=ctx, precision=code)
        f.Analyse()
        out[name] = (ens.data.copy(), f.local_counts())
    prob = orc.Problem(k, case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], f.inflation, f.localization_radius)
    Xo, _, _ = orc.letkf_standard(prob, case["X"], hx)
    np.testing.assert_array_equal(out["stencil"][1], out["hx"][1])
    for name in ("stencil", "hx"):
        Xm = out[name][0]
        rms = np.sqrt(np.mean((Xm - Xo) ** 2)) / np.sqrt(np.mean(Xo ** 2))
        assert rms < bound, (name, rms)


@pytest.mark.parametrize("prec,k", [("fp64", 40), ("bf16x3", 40), ("fp64", 64)])
def test_letkf_column_passes_match_single_pass(mda, ctx, orc, prec, k, monkeypatch):
    """Gram -> solve -> transform over passes of 333 columns (the last one
    short) reproduces the single-pass analysis bit for bit: every column is
    computed from the same inputs whatever pass it falls in (k = 64: the
    eigenvector kernel applies the transform at the pass's column offset)."""
    code = {"fp64": mda._lib.PREC_FP64, "bf16x3": mda._lib.PREC_BF16X3}[prec]
    case = latlon_case(k=k, p=4000, seed=91, cluster=True)
    out = []
    for chunk in (None, "333"):
        if chunk is None:
            monkeypatch.delenv("MDA_LETKF_CHUNK_COLS", raising=False)
        else:
            monkeypatch.setenv("MDA_LETKF_CHUNK_COLS", chunk)
        ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
        obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                              coord=mda._lib.COORD_GEOGRAPHIC)
        geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
        f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                       "localization_radius": 1500.0}),
                      geometry=geom, ctx=ctx, precision=code)
        diag = f.Analyse()
        out.append((ens.data.copy(), diag))
    monkeypatch.delenv("MDA_LETKF_CHUNK_COLS", raising=False)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for key in ("total_local_obs", "max_local_obs", "n_empty_columns", "n_fallback_columns"):
        assert out[0][1][key] == out[1][1][key]
    assert out[1][1]["ms_gram"] > 0 and out[1][1]["ms_transform"] > 0


@pytest.fixture(scope="module")
def k128_case(orc):
    """k = 128 on a 32 x 16 x 3 grid: local counts 97..152, so about half the
    local Gram matrices are rank-deficient; the oracle result is shared."""
    case = latlon_case(nlon=32, nlat=16, nlev=3, k=128, p=3200, seed=128, nmodes=96)
    prob = orc.Problem(128, case["nlon"], case["nlat"], case["nlev"], orc.COORD_GEOGRAPHIC,
                       case["col_lat"], case["col_lon"], case["lat"], case["lon"], case["yo"],
                       case["var"], 1.0499999523162842, 2500.0)
    Xo, _, nloc = orc.letkf_standard(prob, case["X"], case["H"].apply_host(case["X"]))
    return case, Xo, nloc


@pytest.mark.parametrize("method", ["auto", "jacobi"])
def test_letkf_k128_matches_oracle(mda, ctx, k128_case, method, monkeypatch):
    """k = 128 (two wavefronts per column, bisection eigenvalues; or the
    Jacobi kernel in global scratch) against the oracle, run in passes of
    200 columns."""
    case, Xo, nloc = k128_case
    assert nloc.min() < 127 < nloc.max()
    monkeypatch.setenv("MDA_LETKF_CHUNK_COLS", "200")
    ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                          coord=mda._lib.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    em = mda._lib.EIG_AUTO if method == "auto" else mda._lib.EIG_JACOBI
    f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                   "localization_radius": 2500.0}),
                  geometry=geom, ctx=ctx, eig_method=em)
    assert f.inflation == 1.0499999523162842
    diag = f.Analyse()
    monkeypatch.delenv("MDA_LETKF_CHUNK_COLS", raising=False)
    assert diag["total_local_obs"] == nloc.sum()
    assert rel_err(ens.data, Xo) < TOL
    if method == "auto":
        assert diag["n_fallback_columns"] == 0
        assert diag["ms_eig_values"] > 0


@pytest.mark.parametrize("prec,bound", [("fp32", 1e-6), ("bf16x3", 1e-4)])
def test_letkf_k128_mixed_precision(mda, ctx, k128_case, prec, bound, monkeypatch):
    """k = 128 with the FP32 / BF16x3 Gram (two members per lane, one wave
    per column): relative RMS against the oracle's FP64 analysis within
    `bound`, same local counts."""
    case, Xo, nloc = k128_case
    monkeypatch.setenv("MDA_LETKF_CHUNK_COLS", "300")
    code = {"fp32": mda._lib.PREC_FP32, "bf16x3": mda._lib.PREC_BF16X3}[prec]
    ens = mda.Ensemble(case["X"].copy(), case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(case["yo"], case["var"], case["lat"], case["lon"],
                          coord=mda._lib.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, case["col_lat"], case["col_lon"])
    f = mda.LETKF(ens, obs, case["H"], mda.Config({"inflation": 1.05,
                                                   "localization_radius": 2500.0}),
                  geometry=geom, ctx=ctx, precision=code)
    diag = f.Analyse()
    monkeypatch.delenv("MDA_LETKF_CHUNK_COLS", raising=False)
    assert diag["total_local_obs"] == nloc.sum()
    Xm = ens.data
    rms = np.sqrt(np.mean((Xm - Xo) ** 2)) / np.sqrt(np.mean(Xo ** 2))
    assert rms < bound, rms
    inc = np.sqrt(np.mean((Xo - case["X"]) ** 2))
    assert np.sqrt(np.mean((Xm - Xo) ** 2)) < 10 * bound * inc
