"""The forecast step for cycling: Model<Tag> over the Lorenz-96 backend
(Model.hpp:57-177, ModelConcepts.hpp:51-71, lorenz95.py:3-37) through the
Python mirror and through the C++ adapter, against the oracle's Euler and
RK4 steps, and a forecast -> LETKF -> forecast cycle against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from cases import l96_case, rel_err
from test_abi_cpu import build_adapter_mock

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mda():
    import metada_amd
    return metada_amd


@pytest.fixture(scope="module")
def ctx(mda):
    c = mda.Context(0)
    yield c
    c.close()


def _oracle_forecast(orc, X, F, dt, steps, scheme):
    out = np.array(X, dtype=float, copy=True)
    for j in range(out.shape[0]):
        for _ in range(steps):
            out[j] = orc.l96_step(out[j], F, dt, scheme)
    return out


@pytest.mark.parametrize("scheme,steps", [("euler", 1), ("euler", 6), ("rk4", 1), ("rk4", 6)])
def test_model_run_matches_oracle(mda, ctx, orc, scheme, steps):
    rng = np.random.default_rng(steps)
    N = 60
    x0 = 8.0 + rng.standard_normal(N)
    cfg = mda.Config({"scheme": scheme, "num_steps": steps, "dt": 0.02, "forcing": 8.0})
    model = mda.Model(cfg, ctx=ctx)
    xf = np.zeros(N)
    model.run(x0, xf)
    want = _oracle_forecast(orc, x0[None, :], 8.0, float(np.float32(0.02)), steps, scheme)[0]
    assert rel_err(xf, want) < 1e-12 * steps
    # run(initial, initial) advances in place
    x1 = x0.copy()
    model.run(x1, x1)
    np.testing.assert_array_equal(x1, xf)


def test_run_ensemble_on_device_in_one_launch(mda, ctx, orc):
    import torch
    rng = np.random.default_rng(4)
    N, k = 400, 24
    X = 8.0 + rng.standard_normal((k, N))
    model = mda.Lorenz96Model(mda.Config({"scheme": "rk4", "num_steps": 5}), ctx=ctx)
    ens = mda.Ensemble(torch.from_numpy(X.copy()).to("cuda:0"), N, 1)
    model.runEnsemble(ens)
    got = ens.to_host()
    # one member at a time through run() gives the same bits
    for j in (0, 7, k - 1):
        xf = np.zeros(N)
        model.run(X[j].copy(), xf)
        np.testing.assert_array_equal(got[j], xf)
    want = _oracle_forecast(orc, X, 8.0, 0.05, 5, "rk4")
    assert rel_err(got, want) < 1e-11


def test_cpp_model_backend_behind_model_wrapper(ctx, orc, tmp_path):
    exe = build_adapter_mock(tmp_path)
    rng = np.random.default_rng(9)
    N, k, steps = 40, 5, 4
    X = 8.0 + rng.standard_normal((k, N))
    for scheme in ("euler", "rk4"):
        inp = tmp_path / f"l96_{scheme}.in"
        with open(inp, "w") as f:
            f.write(f"l96 {scheme} {N} {k} {steps} 0.03 8.0\n")
            f.write("\n".join(repr(float(v)) for v in X.ravel()) + "\n")
        out = str(tmp_path / f"l96_{scheme}")
        r = subprocess.run([exe, str(inp), out], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        single = np.loadtxt(out + ".single").reshape(k, N)
        ensemble = np.loadtxt(out + ".ensemble").reshape(k, N)
        want = _oracle_forecast(orc, X, 8.0, float(np.float32(0.03)), steps, scheme)
        assert rel_err(single, want) < 1e-12 * steps
        np.testing.assert_array_equal(single, ensemble)
        assert "bad_scheme Lorenz96Model: unknown scheme 'leapfrog'" in r.stdout
        assert "initialized_after_finalize 0" in r.stdout


def test_forecast_analysis_cycle_matches_oracle(mda, ctx, orc):
    """Three cycles of Model.runEnsemble then LETKF.Analyse on Lorenz-96 (N=40,
    k=20, every other point observed through the IDW4 operator), against the
    oracle's RK4 step and LETKF."""
    import torch
    case = l96_case(orc, N=40, k=20, seed=31)
    N, k, p = case["N"], case["k"], case["p"]
    model = mda.Model(mda.Config({"scheme": "rk4", "num_steps": 2}), ctx=ctx)
    ens = mda.Ensemble(torch.from_numpy(case["X"].copy()).to("cuda:0"), N, 1)
    H = mda.ObsOperator(case["rp"], case["cols"], case["ws"])
    cfg = mda.Config({"inflation": 1.1, "localization_radius": 6.0})
    Xo = case["X"].copy()
    truth = case["truth"].copy()
    rng = np.random.default_rng(5)
    for cycle in range(3):
        truth = _oracle_forecast(orc, truth[None, :], 8.0, 0.05, 2, "rk4")[0]
        yo = orc.csr_apply(case["rp"], case["cols"], case["ws"], truth) + rng.standard_normal(p)
        model.backend().runEnsemble(ens)
        Xo = _oracle_forecast(orc, Xo, 8.0, 0.05, 2, "rk4")
        ob = mda.Observation(yo, case["var"], case["obs_x"], case["obs_y"], np.zeros(p))
        f = mda.LETKF(ens, ob, H, cfg, ctx=ctx)
        f.Analyse()
        prob = orc.Problem(k, N, 1, 1, orc.COORD_GRID, np.arange(N, dtype=float), np.zeros(N),
                           case["obs_x"], case["obs_y"], yo, case["var"], f.inflation,
                           f.localization_radius)
        Yb = np.stack([orc.csr_apply(case["rp"], case["cols"], case["ws"], Xo[j])
                       for j in range(k)])
        Xo, _, _ = orc.letkf_standard(prob, Xo, Yb)
        assert rel_err(ens.to_host(), Xo) < 1e-9, cycle
