"""GPU tests of the batching form: independent LETKF analyses of different
ensemble sizes over clustered observations, run as one batch (concurrent
branches of one hipGraph, work-queue Gram, autotuned launch), against the
same analyses run one at a time; graph replays against the eager result."""
import numpy as np
import pytest

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


def _workload(mda, k, seed, p=20000, nlon=96, nlat=48, nlev=4):
    from metada_amd import synthetic
    return synthetic.latlon_workload(nlon, nlat, nlev, k, p, seed=seed, clustered=True,
                                     nblobs=6, blob_km=600.0, device="cuda:0")


def _filter(mda, ctx, W, nlon, nlat, nlev, prec):
    ens = mda.Ensemble(W["X"], nlon, nlat, nlev)
    cfg = mda.Config({"inflation": 1.05, "localization_radius": 1500.0})
    return mda.LETKF(ens, W["obs"], W["H"], cfg, geometry=W["geom"], ctx=ctx, precision=prec)


def test_batch_matches_single_analyses_and_graph_replays(mda, ctx):
    import torch
    L = mda._lib
    specs = [(32, L.PREC_BF16X3, 1), (64, L.PREC_FP64, 2), (20, L.PREC_FP32, 3)]
    work = [(_workload(mda, k, seed), k, prec) for k, prec, seed in specs]
    Xb = [W["X"].clone() for W, _, _ in work]
    # one at a time
    single = []
    for (W, k, prec), xb in zip(work, Xb):
        f = _filter(mda, ctx, W, 96, 48, 4, prec)
        d = f.Analyse()
        single.append((W["X"].clone(), d))
        W["X"].copy_(xb)
    # clustered observations: local counts vary widely, some columns empty
    counts = [d["max_local_obs"] for _, d in single]
    assert all(c > 0 for c in counts)
    filters = [_filter(mda, ctx, W, 96, 48, 4, prec) for W, k, prec in work]
    # graph + work queue (default launch), then autotuned: each call restores
    # the backgrounds and must reproduce the one-at-a-time analyses exactly
    runs = [(L.BATCH_GRAPH | L.BATCH_QUEUE, 0), (L.BATCH_GRAPH | L.BATCH_QUEUE, 1),
            (L.BATCH_GRAPH | L.BATCH_QUEUE, 1),
            (L.BATCH_GRAPH | L.BATCH_QUEUE | L.BATCH_AUTOTUNE, None)]
    for flags, replay in runs:
        for (W, _, _), xb in zip(work, Xb):
            W["X"].copy_(xb)
        torch.cuda.synchronize()
        diags, info = mda.letkf_batch(filters, flags)
        if replay is not None:
            assert info["graph_replayed"] == replay
            assert all(q == 1 for q in info["gram_queue"])
        for (W, _, _), (Xs, ds), d in zip(work, single, diags):
            # per-column arithmetic is the same whatever the processing order
            assert torch.equal(W["X"], Xs)
            assert d["total_local_obs"] == ds["total_local_obs"]
            assert d["n_fallback_columns"] == ds["n_fallback_columns"]
        # eager work between replays
        junk = torch.randn(1 << 20, device="cuda:0").sum().item()
        assert np.isfinite(junk)


def test_batch_without_graph_or_queue(mda, ctx):
    import torch
    L = mda._lib
    W = _workload(mda, 16, 9)
    xb = W["X"].clone()
    f = _filter(mda, ctx, W, 96, 48, 4, L.PREC_FP64)
    f.Analyse()
    ref = W["X"].clone()
    W["X"].copy_(xb)
    diags, info = mda.letkf_batch([_filter(mda, ctx, W, 96, 48, 4, L.PREC_FP64)], 0)
    assert info["graph_replayed"] == 0 and info["gram_queue"] == [0]
    assert torch.equal(W["X"], ref)


def test_batch_rejects_host_inputs(mda, ctx):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((8, 40))
    obs = mda.Observation(rng.standard_normal(5), np.ones(5), np.arange(5.0), np.zeros(5),
                          np.zeros(5))
    H = mda.ObsOperator.identity_idw4(obs, 40, 1, 1)
    f = mda.LETKF(mda.Ensemble(X, 40, 1, 1), obs, H,
                  mda.Config({"inflation": 1.0, "localization_radius": 3.0}), ctx=ctx)
    with pytest.raises(mda._lib.MdaError, match="device-resident"):
        mda.letkf_batch([f])
