"""The allocation-free ETKF (mda_etkf_analyse_noalloc) through the C-ABI
against mda_etkf_analyse on the same device inputs, on two calls in a row
(the second smaller), plus its limits and the oracle's ETKF."""
import ctypes as C

import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu


def _problem(seed, k, n, p):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((k, n)) + np.sin(np.linspace(0, 6, n))
    idx = np.sort(rng.choice(n, p, replace=False)).astype(np.int64)
    rp = np.arange(p + 1, dtype=np.int64)
    w = np.ones(p)
    yo = X.mean(0)[idx] + 0.3 * rng.standard_normal(p)
    var = rng.uniform(0.2, 1.5, p)
    return X, idx, rp, w, yo, var


def _device_structs(mda, X, idx, rp, w, yo, var):
    import torch
    dev = "cuda:0"
    t = {name: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
         for name, a in (("X", X), ("idx", idx), ("rp", rp), ("w", w), ("yo", yo),
                         ("var", var), ("oa", idx.astype(float)), ("ob", np.zeros(len(idx))))}
    ens = mda.Ensemble(t["X"], X.shape[1], 1, 1)
    obs = mda.Observation(t["yo"], t["var"], t["oa"], t["ob"])
    H = mda.ObsOperator(t["rp"], t["idx"], t["w"])
    return t, ens, obs, H


def _call(mda, ctx, fn, ens, obs, H, rho):
    from metada_amd.analysis import _obs_struct
    es, ek = ens.struct()
    os_, ok = _obs_struct(obs)
    hs, hk = H.struct()
    return fn(ctx.h, rho, C.byref(es), C.byref(os_), C.byref(hs))


def test_noalloc_etkf_matches_etkf_on_two_calls():
    import torch
    import metada_amd as mda
    from oracle import oracle as orc
    ref_ctx = mda.Context(0)
    na_ctx = mda.Context(0)  # used only by the allocation-free entry point
    rho = float(np.float32(1.08))
    for seed, (k, n, p) in ((1, (40, 3000, 1200)), (2, (12, 700, 150))):
        print(f"[no_alloc] seed {seed}: k={k} n={n} p={p}", flush=True)
        X, idx, rp, w, yo, var = _problem(seed, k, n, p)
        _, ens_a, obs_a, H_a = _device_structs(mda, X, idx, rp, w, yo, var)
        tb, ens_b, obs_b, H_b = _device_structs(mda, X, idx, rp, w, yo, var)
        ref_ctx.check(_call(mda, ref_ctx, ref_ctx.lib.mda_etkf_analyse, ens_a, obs_a, H_a, rho))
        na_ctx.check(_call(mda, na_ctx, na_ctx.lib.mda_etkf_analyse_noalloc, ens_b, obs_b, H_b,
                           rho))
        torch.cuda.synchronize()
        Xa, Xb = ens_a.data.cpu().numpy(), ens_b.data.cpu().numpy()
        assert rel_err(Xb, Xa) < 1e-10
        np.testing.assert_array_equal(Xb, Xa)  # same kernels, same launch shapes
        Xo = orc.etkf_analyse(X, X[:, idx], yo, var, rho)
        assert rel_err(Xb, Xo) < 1e-10
        # the context the allocation-free calls ran on holds no workspace
        assert na_ctx.workspace_bytes() == 0
    ref_ctx.close()
    na_ctx.close()


def test_noalloc_etkf_limits():
    import metada_amd as mda
    from metada_amd._lib import MdaError, MDA_E_ARG
    ctx = mda.Context(0)
    # k above the limit
    X, idx, rp, w, yo, var = _problem(3, 65, 400, 50)
    _, ens, obs, H = _device_structs(mda, X, idx, rp, w, yo, var)
    rc = _call(mda, ctx, ctx.lib.mda_etkf_analyse_noalloc, ens, obs, H, 1.0)
    assert rc == MDA_E_ARG and b"k <= 64" in ctx.lib.mda_last_error(ctx.h)
    # p (k + 1) above the scratch limit
    X, idx, rp, w, yo, var = _problem(4, 64, 40000, 33000)
    t, ens, obs, H = _device_structs(mda, X, idx, rp, w, yo, var)
    rc = _call(mda, ctx, ctx.lib.mda_etkf_analyse_noalloc, ens, obs, H, 1.0)
    assert rc == MDA_E_ARG and b"scratch limit" in ctx.lib.mda_last_error(ctx.h)
    np.testing.assert_array_equal(t["X"].cpu().numpy(), X)
    # host inputs would need staging buffers
    X, idx, rp, w, yo, var = _problem(5, 8, 100, 10)
    ens = mda.Ensemble(X.copy(), 100, 1, 1)
    obs = mda.Observation(yo, var, idx.astype(float), np.zeros(10))
    H = mda.ObsOperator(rp, idx, w)
    rc = _call(mda, ctx, ctx.lib.mda_etkf_analyse_noalloc, ens, obs, H, 1.0)
    assert rc == MDA_E_ARG
    with pytest.raises(MdaError):
        ctx.check(rc)
    ctx.close()
