"""LWEnKF on the GPU (mda_lwenkf_analyse) against the oracle's restatement of
LWEnKF.hpp:207-334: the reference's tutorial configuration
(src/backends/simple/tutorial/lwenkf.yaml), every localisation function and
weighting scheme, both Gaspari-Cohn forms, k = 128, a device-resident
ensemble, the C++ adapter, and the refused inputs."""
import os
import subprocess

import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUT = os.path.join(ROOT, "tests", "golden", "tutorial")


@pytest.fixture(scope="module")
def mda():
    import metada_amd
    return metada_amd


@pytest.fixture(scope="module")
def ctx(mda):
    c = mda.Context(0)
    yield c
    c.close()


def _cfg(mda, infl, radius, fn, weighting, method="multiplicative"):
    return mda.Config({"analysis": {"inflation": infl, "localization_radius": radius,
                                    "inflation_method": method, "localization_function": fn,
                                    "weighting_scheme": weighting}})


def _ring_case(orc, k, n, p, seed, var_scale=1.0):
    """Members of a smooth 1-D field on n points, p observations at random
    positions through the IDW4 operator, y = truth + noise."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 2 * np.pi, n, endpoint=False)
    truth = np.sin(3 * x) + 0.5 * np.cos(7 * x)
    X = truth + 0.6 * rng.standard_normal((k, 1)) * np.sin(x + rng.uniform(0, 6, (k, 1))) \
        + 0.3 * rng.standard_normal((k, n))
    ox = np.sort(rng.uniform(0, n - 1, p))
    var = var_scale * rng.uniform(0.5, 1.5, p)
    rp, cols, ws = [0], [], []
    for xo in ox:
        idx, w = orc.idw4_stencil(n, 1, 1, float(xo), 0.0, 0.0)
        cols += list(idx)
        ws += list(w / w.sum())
        rp.append(len(cols))
    rp, cols, ws = np.array(rp), np.array(cols), np.array(ws)
    yo = orc.csr_apply(rp, cols, ws, truth) + np.sqrt(var) * rng.standard_normal(p)
    Yb = np.stack([orc.csr_apply(rp, cols, ws, X[j]) for j in range(k)])
    Z = rng.standard_normal((p, k))
    return dict(X=X, Yb=Yb, yo=yo, var=var, Z=Z, rp=rp, cols=cols, ws=ws, ox=ox)


def _run(mda, ctx, case, n, cfg, gc_form="reference", device=False):
    X = case["X"].copy()
    if device:
        import torch
        X = torch.from_numpy(X).to("cuda:0")
    ens = mda.Ensemble(X, n, 1, 1)
    p = len(case["yo"])
    ob = mda.Observation(case["yo"], case["var"], case["ox"], np.zeros(p), np.zeros(p))
    H = mda.ObsOperator(case["rp"], case["cols"], case["ws"])
    f = mda.LWEnKF(ens, ob, H, cfg, ctx=ctx, gc_form=gc_form)
    f.Analyse(Z=case["Z"])
    return ens.to_host(), f


def _check_diag(got, want, tol=1e-9):
    for key in ("innovation_norm", "background_spread", "analysis_spread", "max_kalman_gain",
                "min_kalman_gain", "condition_number", "max_weight", "min_weight",
                "weight_variance"):
        w = want[key]
        assert abs(got[key] - w) <= tol * max(abs(w), 1e-300) + 1e-300, (key, got[key], w)


def test_lwenkf_tutorial_matches_oracle(mda, ctx, orc):
    """lwenkf.yaml: ens_1..8 on the 36 x 18 grid, the 28 observations of
    obs_format.txt (error 0.1), inflation 1.1, radius 0.5, gaussian taper,
    adaptive weights."""
    X = np.stack([mda.read_simple_state(os.path.join(TUT, f"ens_{i}.txt")) for i in range(1, 9)])
    err = mda.Config({"e": 0.1}).Get("e").asFloat()
    obs = mda.Observation.from_grid_file(os.path.join(TUT, "obs_format.txt"), error=err)
    H = mda.ObsOperator.identity_idw4(obs, 36, 18, 1, ctx=ctx)
    k, p = X.shape[0], obs.size()
    Z = np.random.default_rng(2).standard_normal((p, k))
    ens = mda.Ensemble(X.copy(), 36, 18, 1)
    f = mda.LWEnKF(ens, obs, H, _cfg(mda, 1.1, 0.5, "gaussian", "adaptive"), ctx=ctx)
    f.Analyse(Z=Z)
    Yb = np.array([[orc.idw4(X[j], 36, 18, 1, obs.loc_a[o], obs.loc_b[o], obs.loc_c[o])
                    for o in range(p)] for j in range(k)])
    Xo, w, d, _ = orc.lwenkf_analyse(X, Yb, obs.values, obs.variances, Z, f.inflation_factor,
                                     f.localization_radius, "gaussian", "adaptive")
    assert rel_err(ens.to_host(), Xo) < 1e-10
    np.testing.assert_allclose(f.weights, w, rtol=1e-12)
    r = f.getAnalysisResults()
    _check_diag(r, d)
    assert r["localization_function"] == "gaussian" and r["weighting_scheme"] == "adaptive"
    assert r["inflation_method"] == "multiplicative" and r["analysis_increment_norm"] == 0.0
    assert r["ensemble_size"] == 8 and r["observation_count"] == 28


@pytest.mark.parametrize("fn,gc", [("gaussian", "reference"), ("exponential", "reference"),
                                   ("cutoff", "reference"), ("gaspari_cohn", "reference"),
                                   ("gaspari_cohn", "standard")])
@pytest.mark.parametrize("weighting", ["uniform", "adaptive", "inverse_var", "likelihood"])
def test_lwenkf_forms_match_oracle(mda, ctx, orc, fn, gc, weighting):
    k, n, p = 20, 150, 45
    case = _ring_case(orc, k, n, p, seed=len(fn) * 7 + len(weighting), var_scale=0.5)
    cfg = _cfg(mda, 1.05, 0.3, fn, weighting)
    Xg, f = _run(mda, ctx, case, n, cfg, gc_form=gc)
    Xo, w, d, _ = orc.lwenkf_analyse(case["X"], case["Yb"], case["yo"], case["var"], case["Z"],
                                     f.inflation_factor, f.localization_radius, fn, weighting,
                                     gc_standard=(gc == "standard"))
    assert rel_err(Xg, Xo) < 1e-10
    np.testing.assert_allclose(f.weights, w, rtol=1e-11)
    _check_diag(f.getAnalysisResults(), d, tol=1e-8)


@pytest.mark.parametrize("k,n,p", [(64, 6000, 500), (128, 3000, 200)])
def test_lwenkf_large_matches_oracle(mda, ctx, orc, k, n, p):
    case = _ring_case(orc, k, n, p, seed=k)
    cfg = _cfg(mda, 1.1, 0.05, "gaspari_cohn", "inverse_var")
    Xg, f = _run(mda, ctx, case, n, cfg, gc_form="standard", device=True)
    Xo, w, d, _ = orc.lwenkf_analyse(case["X"], case["Yb"], case["yo"], case["var"], case["Z"],
                                     f.inflation_factor, f.localization_radius, "gaspari_cohn",
                                     "inverse_var", gc_standard=True)
    assert rel_err(Xg, Xo) < 1e-10
    _check_diag(f.getAnalysisResults(), d, tol=1e-8)


def test_lwenkf_device_ensemble_equals_host(mda, ctx, orc):
    case = _ring_case(orc, 16, 300, 60, seed=4)
    cfg = _cfg(mda, 1.2, 0.2, "exponential", "likelihood", method="relaxation")
    a, fa = _run(mda, ctx, case, 300, cfg)
    b, fb = _run(mda, ctx, case, 300, cfg, device=True)
    np.testing.assert_array_equal(a, b)
    assert fa.getAnalysisResults()["inflation_method"] == "relaxation"


def test_lwenkf_cpp_adapter(mda, ctx, orc, tmp_path):
    from test_abi_cpu import build_adapter_mock
    exe = build_adapter_mock(tmp_path)
    n, k, p = 40, 10, 12
    case = _ring_case(orc, k, n, p, seed=9)
    ox = np.round(case["ox"]).astype(int)
    Yb = np.array([[orc.idw4(case["X"][j], n, 1, 1, float(x), 0.0, 0.0) for x in ox]
                   for j in range(k)])
    errs = np.sqrt(case["var"])
    inp = tmp_path / "lw.in"
    with open(inp, "w") as fh:
        fh.write(f"lwenkf cutoff:adaptive {n} 1 1 {k} {p} 1.1 0.4\n")
        fh.write("\n".join(repr(float(v)) for v in case["X"].ravel()) + "\n")
        for x, v, e in zip(ox, case["yo"], errs):
            fh.write(f"{int(x)} 0 0 {float(v)!r} {float(e)!r}\n")
        fh.write("\n".join(repr(float(v)) for v in case["Z"].ravel()) + "\n")
    out = str(tmp_path / "lw")
    r = subprocess.run([exe, str(inp), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    diag = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    Xa = np.loadtxt(out + ".analysis").reshape(k, n)
    rho, rad = float(np.float32(1.1)), float(np.float32(0.4))
    Xo, w, d, _ = orc.lwenkf_analyse(case["X"], Yb, case["yo"], errs * errs, case["Z"], rho, rad,
                                     "cutoff", "adaptive")
    assert rel_err(Xa, Xo) < 1e-10
    assert abs(float(diag["condition_number"]) - d["condition_number"]) < 1e-8 * d["condition_number"]
    assert diag["localization_function"] == "cutoff" and diag["weighting_scheme"] == "adaptive"
    assert int(diag["perturbations"]) == k  # RecomputePerturbations (LWEnKF.hpp:304)
    np.testing.assert_allclose(np.loadtxt(out + "_saved_mean.txt").ravel(), Xo.mean(0),
                               atol=6e-7)


def test_lwenkf_refused_inputs(mda, ctx, orc):
    L = mda._lib
    case = _ring_case(orc, 8, 50, 10, seed=1)
    cfg = _cfg(mda, 1.0, 0.5, "gaussian", "uniform")
    bad = dict(case)
    bad["var"] = case["var"].copy()
    bad["var"][3] = np.inf
    with pytest.raises(mda._lib.MdaError) as e:
        _run(mda, ctx, bad, 50, cfg)
    assert e.value.code == L.MDA_E_ARG and "invalid observations" in str(e.value)
    far = dict(case)
    far["yo"] = case["yo"] + 1e4
    with pytest.raises(mda._lib.MdaError) as e:
        _run(mda, ctx, far, 50, _cfg(mda, 1.0, 0.5, "gaussian", "likelihood"))
    assert e.value.code == L.MDA_E_NUMERIC
    nan = dict(case)
    nan["yo"] = case["yo"].copy()
    nan["yo"][0] = np.nan
    with pytest.raises(mda._lib.MdaError) as e:
        _run(mda, ctx, nan, 50, cfg)
    assert e.value.code == L.MDA_E_NUMERIC
    one = dict(case)
    one["X"] = case["X"][:1]
    one["Z"] = case["Z"][:, :1]
    with pytest.raises(mda._lib.MdaError) as e:
        _run(mda, ctx, one, 50, cfg)
    assert e.value.code == L.MDA_E_ARG
    # the ensemble is untouched by a refused call
    X = case["X"].copy()
    ens = mda.Ensemble(X, 50, 1, 1)
    ob = mda.Observation(bad["yo"], bad["var"], case["ox"], np.zeros(10), np.zeros(10))
    f = mda.LWEnKF(ens, ob, mda.ObsOperator(case["rp"], case["cols"], case["ws"]), cfg, ctx=ctx)
    with pytest.raises(mda._lib.MdaError):
        f.Analyse(Z=case["Z"])
    np.testing.assert_array_equal(X, case["X"])
