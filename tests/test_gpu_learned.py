"""GPU tests of the learned-inflation path: the Lorenz-96 forecast kernel and
the MLP kernels against the CPU oracle, the local statistics against a
brute-force restatement, and the twin-experiment protocol end to end on a
small configuration."""
import ctypes as C

import numpy as np
import pytest

from cases import rel_err

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


@pytest.mark.parametrize("N,scheme", [(40, "rk4"), (4000, "rk4"), (40, "euler"), (1000, "euler")])
def test_l96_forecast_matches_oracle(mda, ctx, orc, N, scheme):
    rng = np.random.default_rng(N)
    X = 8.0 + rng.standard_normal((5, N))
    for nsteps in (1, 10):
        ref = X.copy()
        for j in range(5):
            for _ in range(nsteps):
                ref[j] = orc.l96_step(ref[j], 8.0, 0.05 if scheme == "rk4" else 0.01, scheme)
        out = X.copy()
        ctx.check(ctx.lib.mda_l96_forecast(ctx.h, out.ctypes.data, 5, N, 8.0,
                                           0.05 if scheme == "rk4" else 0.01, nsteps,
                                           0 if scheme == "rk4" else 1, 0))
        assert rel_err(out, ref) < 1e-12 * nsteps


def _mlp_inputs(orc, seed, m):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(orc.MLP_NPARAM) * 0.3).astype(np.float32)
    feat = rng.standard_normal((m, 8)).astype(np.float32)
    target = (1.0 + rng.random(m)).astype(np.float32)
    return p, feat, target


@pytest.mark.parametrize("m", [1, 300, 5000])
def test_mlp_forward_and_gradient_match_oracle(mda, ctx, orc, m):
    """FP32 kernels against the FP64 oracle on the same (FP32) inputs."""
    p, feat, target = _mlp_inputs(orc, 3 + m, m)
    rho = np.empty(m, np.float32)
    ctx.check(ctx.lib.mda_mlp_forward(ctx.h, p.ctypes.data, feat.ctypes.data, m,
                                      rho.ctypes.data, 0))
    rho_o = orc.mlp_forward(p.astype(np.float64), feat.astype(np.float64))
    np.testing.assert_allclose(rho, rho_o, rtol=2e-5)
    loss = C.c_double(0)
    grad = np.empty(orc.MLP_NPARAM, np.float32)
    ctx.check(ctx.lib.mda_mlp_loss_grad(ctx.h, p.ctypes.data, feat.ctypes.data,
                                        target.ctypes.data, m, C.byref(loss), grad.ctypes.data, 0))
    lo, go = orc.mlp_loss_grad(p.astype(np.float64), feat.astype(np.float64),
                               target.astype(np.float64))
    assert loss.value == pytest.approx(lo, rel=1e-4)
    assert np.linalg.norm(grad - go) <= 1e-4 * np.linalg.norm(go)
    np.testing.assert_allclose(grad, go, rtol=1e-3, atol=1e-5 * np.abs(go).max())


def test_mlp_adam_step_matches_numpy(mda, ctx, orc):
    import torch
    rng = np.random.default_rng(5)
    p = rng.standard_normal(orc.MLP_NPARAM).astype(np.float32)
    m1 = (rng.standard_normal(orc.MLP_NPARAM) * 0.1).astype(np.float32)
    m2 = (rng.random(orc.MLP_NPARAM) * 0.01).astype(np.float32)
    g = rng.standard_normal(orc.MLP_NPARAM).astype(np.float32)
    step, lr, b1, b2, eps = 7, 1e-3, 0.9, 0.999, 1e-8
    dev = [torch.from_numpy(v.copy()).cuda() for v in (p, m1, m2, g)]
    ctx.check(ctx.lib.mda_mlp_adam_step(ctx.h, dev[0].data_ptr(), dev[1].data_ptr(),
                                        dev[2].data_ptr(), dev[3].data_ptr(), step, lr, b1, b2,
                                        eps, 1))
    a = b1 * m1.astype(np.float64) + (1 - b1) * g
    v = b2 * m2.astype(np.float64) + (1 - b2) * g.astype(np.float64) ** 2
    ref = p - lr * (a / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    np.testing.assert_allclose(dev[0].cpu().numpy(), ref, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(dev[1].cpu().numpy(), a, rtol=1e-6, atol=2e-7)  # FP32 rounding of g


def test_inflation_features_match_brute_force(mda, ctx, orc):
    """The eight local statistics against a restatement with the oracle's
    haversine and Gaspari-Cohn over every observation."""
    rng = np.random.default_rng(11)
    N, k, every = 200, 20, 2
    X = 8.0 + rng.standard_normal((k, N)) * (1.0 + 0.5 * np.sin(np.arange(N) / 9.0))
    lon = np.arange(N) * 360.0 / N
    oi = np.arange(0, N, every)
    p = len(oi)
    var = 0.5 + rng.random(p)
    yo = X.mean(0)[oi] + rng.standard_normal(p)
    radius = 12 * 2 * np.pi * 6371.0 / N
    ens = mda.Ensemble(X.copy(), N, 1, 1)
    obs = mda.Observation(yo, var, np.zeros(p), lon[oi], coord=mda._lib.COORD_GEOGRAPHIC)
    H = mda.ObsOperator(np.arange(p + 1, dtype=np.int64), oi.astype(np.int64), np.ones(p))
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, np.zeros(N), lon)
    f = mda.LETKF(ens, obs, H, mda.Config({"inflation": 1.05, "localization_radius": radius}),
                  geometry=geom, ctx=ctx)
    from metada_amd.analysis import _obs_struct
    es, ek = ens.struct()
    gs, gk = geom.struct()
    os_, ok = _obs_struct(obs)
    hs, hk = H.struct()
    feat = np.empty((N, 8), np.float32)
    ctx.check(ctx.lib.mda_letkf_inflation_features(ctx.h, C.byref(f.params), C.byref(es),
                                                   C.byref(gs), C.byref(os_), C.byref(hs), None,
                                                   feat.ctypes.data, 0))
    Yb = X[:, oi]
    Yp = Yb - Yb.mean(0)
    d = yo - Yb.mean(0)
    sp = (Yp ** 2).sum(0) / (k - 1)
    c = float(np.float32(radius)) / 2.0
    ref = np.zeros((N, 8))
    for g in range(N):
        dist = np.array([orc.haversine_km(0.0, lon[g], 0.0, lon[o]) for o in oi])
        sel = dist <= float(np.float32(radius))
        w = np.array([orc.gaspari_cohn(z) for z in dist[sel] / c]).clip(0) / var[sel]
        Sw, Sd2, Sd, Ssp = w.sum(), (w * d[sel] ** 2).sum(), (w * d[sel]).sum(), (w * sp[sel]).sum()
        v = X[:, g].var(ddof=1)
        ref[g] = [np.log(Sd2 / Sw), np.log(Ssp / Sw), np.log(Sd2 / (Ssp + Sw)),
                  Sd / np.sqrt(Sw * Sd2), np.log(Sw), np.log1p(sel.sum()), np.log(v), 0.0]
    np.testing.assert_allclose(feat, ref, rtol=2e-5, atol=2e-6)


def test_learned_inflation_gain_over_mistuned_fixed_rho(mda, ctx):
    """The twin-experiment protocol end to end on L96 (N = 400, k = 20): the
    plain filter with an over-inflating fixed rho = 1.3 gives the training
    data; the network trained on it by the library's MLP kernels is chosen on
    the validation cycles and then beats fixed inflation on later evaluation
    cycles by a wide margin (the fixed filter's spread is far above its
    error; the learned filter's is not)."""
    from metada_amd import learned
    res, = learned.experiment(learned.L96Setup(N=400, k=20, support_gp=10.0, rho_fixed=1.3),
                              train_cycles=200, eval_cycles=300, val_cycles=60, burn_in=100,
                              train_steps=400, ctx=ctx)
    assert res["train_loss_last"] < res["train_loss_first"]
    assert res["train_diverged_at"] is None and res["learned_diverged_at"] is None
    assert res["rmse_fixed"] < 1.0  # obs error is 1: the filter beats the observations
    assert res["rmse_val_learned"] < res["rmse_val_fixed"] and not res["fallback"]
    assert res["rmse_deployed"] == res["rmse_learned"]
    assert res["gain"] > 0.15, res
    assert res["spread_fixed"] > 1.2 * res["rmse_fixed"]


def test_learned_inflation_falls_back_when_not_better(mda, ctx):
    """With a well-tuned fixed rho the choice on the validation cycles decides
    between the two filters, and the deployed evaluation RMSE is that of the
    chosen one (never a choice made on the reported cycles)."""
    from metada_amd import learned
    res, = learned.experiment(learned.L96Setup(N=400, k=20, support_gp=10.0, rho_fixed=1.05),
                              train_cycles=150, eval_cycles=200, val_cycles=60, burn_in=50,
                              train_steps=200, ctx=ctx)
    assert np.isfinite(res["rmse_fixed"])
    chosen = res["rmse_learned"] if not res["fallback"] else res["rmse_fixed"]
    assert res["rmse_deployed"] == chosen
    assert res["fallback"] == (res["rmse_val_learned"] > res["rmse_val_fixed"])
