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


def _device_analysis(mda, ctx, case, prec, radius):
    """A LETKF over device copies of a host test case (latlon_case)."""
    import torch
    dev = "cuda:0"
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    H = case["H"]
    keep = {"X": t(case["X"]), "yo": t(case["yo"]), "var": t(case["var"]), "lat": t(case["lat"]),
            "lon": t(case["lon"]), "rp": t(H.row_ptr), "col": t(H.col), "w": t(H.weight),
            "clat": t(case["col_lat"]), "clon": t(case["col_lon"])}
    ens = mda.Ensemble(keep["X"], case["nlon"], case["nlat"], case["nlev"])
    obs = mda.Observation(keep["yo"], keep["var"], keep["lat"], keep["lon"],
                          coord=mda._lib.COORD_GEOGRAPHIC)
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, keep["clat"], keep["clon"])
    f = mda.LETKF(ens, obs, mda.ObsOperator(keep["rp"], keep["col"], keep["w"]),
                  mda.Config({"inflation": 1.05, "localization_radius": radius}),
                  geometry=geom, ctx=ctx, precision=prec)
    f._keep = keep
    return f


def test_batch_mixed_k_with_k128_matches_oracle(mda, ctx, orc):
    """The batching form on clustered observations with k = 32, 64 and 128
    in one batch (hipGraph, work-queue Gram, autotuned launch): every
    analysis against the oracle's standard LETKF -- FP64 to 1e-10 (max-norm
    relative), BF16x3 to 1e-4 relative RMS -- and the same batch replayed."""
    import torch
    from cases import latlon_case, rel_err
    L = mda._lib
    radius = 1500.0
    specs = [(32, L.PREC_BF16X3), (64, L.PREC_FP64), (128, L.PREC_FP64), (128, L.PREC_BF16X3)]
    cases = [latlon_case(nlon=32, nlat=16, nlev=3, k=k, p=5000, seed=200 + i, cluster=True)
             for i, (k, _) in enumerate(specs)]
    filters = [_device_analysis(mda, ctx, c, prec, radius) for c, (_, prec) in zip(cases, specs)]
    rho = filters[0].inflation
    want = []
    for c in cases:
        prob = orc.Problem(c["k"], c["nlon"], c["nlat"], c["nlev"], orc.COORD_GEOGRAPHIC,
                           c["col_lat"], c["col_lon"], c["lat"], c["lon"], c["yo"], c["var"],
                           rho, radius)
        Xo, _, nloc = orc.letkf_standard(prob, c["X"], c["H"].apply_host(c["X"]))
        want.append((Xo, nloc))
    assert any((nloc == 0).any() for _, nloc in want)  # clustered: empty columns
    flags = L.BATCH_GRAPH | L.BATCH_QUEUE | L.BATCH_AUTOTUNE
    for it in range(2):
        for f, c in zip(filters, cases):
            f._keep["X"].copy_(torch.from_numpy(c["X"]))
        torch.cuda.synchronize()
        diags, info = mda.letkf_batch(filters, flags)
        assert info["graph_replayed"] == it
        for f, (k, prec), (Xo, nloc), d in zip(filters, specs, want, diags):
            Xg = f._keep["X"].cpu().numpy()
            assert d["total_local_obs"] == int(nloc.sum())
            if prec == L.PREC_FP64:
                assert rel_err(Xg, Xo) < 1e-10, (k, rel_err(Xg, Xo))
            else:
                rms = np.sqrt(np.mean((Xg - Xo) ** 2)) / np.sqrt(np.mean(Xo ** 2))
                assert rms < 1e-4, (k, rms)


def test_batch_graph_key_follows_every_input(mda, ctx):
    """A cached graph is replayed only for the same buffers and parameters:
    a new variance array (other values, other address) gives a new graph and
    the single-analysis result."""
    import torch
    from cases import latlon_case
    L = mda._lib
    case = latlon_case(nlon=24, nlat=12, nlev=2, k=16, p=1500, seed=77, cluster=True)
    f = _device_analysis(mda, ctx, case, L.PREC_FP64, 1500.0)
    x0 = torch.from_numpy(case["X"]).to("cuda:0")

    f.Analyse()  # the single analysis's own workspace exists before any graph

    def batch_then_single(filt):
        filt._keep["X"].copy_(x0)
        diags, info = mda.letkf_batch([filt], L.BATCH_GRAPH | L.BATCH_QUEUE)
        got = filt._keep["X"].clone()
        filt._keep["X"].copy_(x0)
        filt.Analyse()
        return info, got, filt._keep["X"].clone()
    info, got, ref = batch_then_single(f)
    assert torch.equal(got, ref)
    info, got, ref = batch_then_single(f)
    assert info["graph_replayed"] == 1 and torch.equal(got, ref)
    # new variances: a new buffer with other values
    f._keep["var"] = f._keep["var"] * 1.7
    f.obs.variances = f._keep["var"]
    info, got, ref = batch_then_single(f)
    assert info["graph_replayed"] == 0 and torch.equal(got, ref)
