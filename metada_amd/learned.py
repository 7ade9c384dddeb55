"""Learned per-column inflation for the LETKF (no counterpart in the reference).

A small MLP (8 -> 64 -> 64 -> 1, tanh, rho = 1 + softplus, FP32) maps eight
local innovation and spread statistics of a column (mda_letkf_inflation_features)
to the prior inflation rho_g of that column's LETKF analysis
(mda_letkf_params.rho_column, Hunt et al. 2007: A = (k-1)/rho_g I + ...).

It is trained on the plain filter's own cycled twin experiment: a cycled
Lorenz-96 LETKF with fixed inflation records, before every analysis, the
features of every column and, as target, the inflation of the current
background that would make its local spread match its local error against
the truth,
    rho*_g = sum_j c_gj (xbar_j - x_t,j)^2 / sum_j c_gj s_j^2,
with c the Gaspari-Cohn weights of the localisation, clipped to [0, 4]
against outliers; the network's rho >= 1 cannot follow targets below 1 and
settles at 1 there.  Runs at several fixed rho give it both under- and
over-dispersive backgrounds.  Training (forward,
backward and Adam) runs in the library's own HIP kernels (mda_mlp_*).  At
evaluation the learned and the fixed-inflation filters are cycled side by
side from the same state; the learned one is kept only if its analysis RMSE
against the truth (Metrics.hpp:265-273) is not worse on a validation
segment, otherwise the result falls back to fixed inflation, and the choice
is then scored on later cycles it has not seen.

Lorenz-96 (tests/python/lorenz95.py:3-37) is stepped by the library's GPU
RK4 kernel (mda_l96_forecast); the N grid points sit on the equator, so the
LETKF's haversine distance is the cyclic index distance times 2 pi R / N.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .analysis import Config, ColumnGeometry, Context, Ensemble, LETKF, ObsOperator, Observation

EARTH_R_KM = 6371.0


def _torch():
    import torch
    return torch


class Inflator:
    """The MLP and its Adam state, all FP32 device tensors; feature
    standardisation (mean, std) is part of the model."""

    def __init__(self, ctx: Context, seed: int = 0, device: str = "cuda:0"):
        torch = _torch()
        self.ctx = ctx
        g = torch.Generator().manual_seed(seed)
        parts = []
        for fan_in, fan_out in ((L.MLP_NIN, L.MLP_HIDDEN), (L.MLP_HIDDEN, L.MLP_HIDDEN)):
            lim = math.sqrt(6.0 / (fan_in + fan_out))
            parts.append((torch.rand(fan_out * fan_in, generator=g) * 2 - 1) * lim)
            parts.append(torch.zeros(fan_out))
        lim = math.sqrt(6.0 / (L.MLP_HIDDEN + 1))
        parts.append((torch.rand(L.MLP_HIDDEN, generator=g) * 2 - 1) * lim)
        parts.append(torch.full((1,), -4.0))  # softplus(-4) = 0.018: rho starts near 1.02
        self.params = torch.cat(parts).float().to(device).contiguous()
        assert self.params.numel() == L.MLP_NPARAM
        self.m1 = torch.zeros_like(self.params)
        self.m2 = torch.zeros_like(self.params)
        self.step = 0
        self.mu = torch.zeros(L.MLP_NIN, device=device)
        self.sd = torch.ones(L.MLP_NIN, device=device)
        self.device = device

    def normalise(self, feat):
        return ((feat - self.mu) / self.sd).contiguous()

    def fit_normalisation(self, feat):
        torch = _torch()
        self.mu = feat.mean(0)
        sd = feat.std(0)
        # a feature constant over the training data (e.g. the previous
        # inflation when one fixed rho made all of it) keeps unit scale, so
        # other values at inference stay of order one
        self.sd = torch.where(sd > 1e-6, sd, torch.ones_like(sd))

    def predict(self, feat):
        torch = _torch()
        x = self.normalise(feat)
        rho = torch.empty(x.shape[0], device=self.device, dtype=torch.float32)
        self.ctx.check(self.ctx.lib.mda_mlp_forward(self.ctx.h, self.params.data_ptr(), x.data_ptr(),
                                                    x.shape[0], rho.data_ptr(), 1))
        return rho

    def loss_grad(self, x, target):
        torch = _torch()
        grad = torch.empty_like(self.params)
        loss = C.c_double(0.0)
        self.ctx.check(self.ctx.lib.mda_mlp_loss_grad(self.ctx.h, self.params.data_ptr(),
                                                      x.data_ptr(), target.data_ptr(),
                                                      x.shape[0], C.byref(loss),
                                                      grad.data_ptr(), 1))
        return loss.value, grad

    def train(self, feat, target, steps=3000, batch=32768, lr=3e-3, seed=0):
        """Adam on mini-batches of (standardised features, target rho)."""
        torch = _torch()
        bad_f = (~torch.isfinite(feat)).sum(0)
        nt = int((~torch.isfinite(target)).sum())
        if int(bad_f.sum()) or nt:
            raise ValueError(f"training data not finite: {bad_f.tolist()} values per feature, "
                             f"{nt} targets")
        self.fit_normalisation(feat)
        x_all = self.normalise(feat)
        t_all = target.float().contiguous()
        g = torch.Generator(device=self.device).manual_seed(seed)
        history = []
        for it in range(steps):
            idx = torch.randint(0, x_all.shape[0], (min(batch, x_all.shape[0]),), generator=g,
                                device=self.device)
            x = x_all[idx].contiguous()
            t = t_all[idx].contiguous()
            loss, grad = self.loss_grad(x, t)
            if not math.isfinite(loss):
                raise FloatingPointError(f"MLP loss is not finite at step {it}")
            self.step += 1
            self.ctx.check(self.ctx.lib.mda_mlp_adam_step(self.ctx.h, self.params.data_ptr(),
                                                          self.m1.data_ptr(), self.m2.data_ptr(),
                                                          grad.data_ptr(), self.step, lr, 0.9,
                                                          0.999, 1e-8, 1))
            history.append(loss)
        return history


@dataclass
class L96Setup:
    N: int = 4000
    k: int = 40
    dt: float = 0.05
    forcing: float = 8.0
    obs_every: int = 2
    obs_sigma: float = 1.0
    support_gp: float = 16.0     # Gaspari-Cohn support in grid points
    rho_fixed: float = 1.05
    seed: int = 2024
    spinup_steps: int = 1000

    @property
    def km_per_gp(self):
        return 2.0 * math.pi * EARTH_R_KM / self.N

    @property
    def radius_km(self):
        return self.support_gp * self.km_per_gp


@dataclass
class CycleLog:
    rmse: list = field(default_factory=list)
    spread: list = field(default_factory=list)
    rho_mean: list = field(default_factory=list)
    diverged_at: int | None = None


class L96Twin:
    """Truth, synthetic observations and one LETKF ensemble of a cycled
    Lorenz-96 twin experiment, device-resident."""

    def __init__(self, ctx: Context, s: L96Setup, device: str = "cuda:0"):
        torch = _torch()
        self.ctx, self.s, self.device = ctx, s, device
        N = s.N
        x = np.full(N, s.forcing)
        x[19 % N] += 0.001  # perturbed rest state, as in the reference's L96 test
        self.truth = torch.tensor(x, device=device).view(1, N).contiguous()
        self.forecast(self.truth, s.spinup_steps)
        self.gen = torch.Generator(device=device).manual_seed(s.seed)
        self.X = (self.truth + torch.randn(s.k, N, generator=self.gen, device=device,
                                           dtype=torch.float64)).contiguous()
        lon = np.arange(N) * 360.0 / N
        self.geom = ColumnGeometry(L.COORD_GEOGRAPHIC, np.zeros(N), lon)
        oi = np.arange(0, N, s.obs_every)
        self.obs_idx = torch.tensor(oi, device=device)
        self.p = len(oi)
        self.H = ObsOperator(np.arange(self.p + 1, dtype=np.int64), oi.astype(np.int64),
                             np.ones(self.p))
        self.obs_lat = np.zeros(self.p)
        self.obs_lon = lon[oi]
        self.var = np.full(self.p, s.obs_sigma ** 2)
        # localisation window weights for the training target (cyclic, GC)
        half = int(math.ceil(s.support_gp))
        d = np.abs(np.arange(-half, half + 1)) / (s.support_gp / 2.0)
        self.gc = torch.tensor([_gc(z) for z in d], device=device, dtype=torch.float64)

    def forecast(self, X, nsteps=1):
        self.ctx.check(self.ctx.lib.mda_l96_forecast(self.ctx.h, X.data_ptr(), X.shape[0],
                                                     X.shape[1], self.s.forcing, self.s.dt,
                                                     nsteps, 0, 1))

    def observe(self):
        torch = _torch()
        y = self.truth[0, self.obs_idx] + self.s.obs_sigma * torch.randn(
            self.p, generator=self.gen, device=self.device, dtype=torch.float64)
        return y

    def obs(self, y):
        return Observation(y.cpu().numpy(), self.var, self.obs_lat, self.obs_lon,
                           coord=L.COORD_GEOGRAPHIC)

    def config(self, rho=None):
        return Config({"inflation": self.s.rho_fixed if rho is None else rho,
                       "localization_radius": self.s.radius_km})

    def target_rho(self, X):
        """rho*_g: the background's error-to-spread ratio against the truth
        around column g (see module docstring), the inflation of the current
        background that would make its spread match its error."""
        torch = _torch()
        mean = X.mean(0)
        err = (mean - self.truth[0]) ** 2
        var = X.var(0, unbiased=True)
        half = (self.gc.numel() - 1) // 2
        ke = self.gc.view(1, 1, -1)
        pad = lambda v: torch.cat([v[-half:], v, v[:half]]).view(1, 1, -1)
        e = torch.nn.functional.conv1d(pad(err), ke).view(-1)
        s = torch.nn.functional.conv1d(pad(var), ke).view(-1)
        # no lower clip at 1: where the background is over-dispersive the
        # target is below 1 and pulls the network's output (>= 1) down to 1
        return (e / s.clamp_min(1e-12)).clamp(0.0, 4.0)

    def features(self, y, cfg):
        """The eight local statistics of every column (mda_letkf_inflation_features)."""
        torch = _torch()
        f = LETKF(Ensemble(self.X, self.s.N, 1, 1), self.obs(y), self.H, cfg,
                  geometry=self.geom, ctx=self.ctx)
        out = torch.empty(self.s.N, L.MLP_NIN, device=self.device, dtype=torch.float32)
        from .analysis import _obs_struct
        es, ek = f.ensemble.struct()
        gs, gk = self.geom.struct()
        os_, ok = _obs_struct(f.obs)
        hs, hk = self.H.struct()
        self.ctx.check(self.ctx.lib.mda_letkf_inflation_features(
            self.ctx.h, C.byref(f.params), C.byref(es), C.byref(gs), C.byref(os_), C.byref(hs),
            None, out.data_ptr(), 1))
        del ek, gk, ok, hk  # staging buffers stay alive until the call returns
        return out

    def analyse(self, y, cfg, rho_col=None):
        f = LETKF(Ensemble(self.X, self.s.N, 1, 1), self.obs(y), self.H, cfg,
                  geometry=self.geom, ctx=self.ctx,
                  rho_column=None if rho_col is None else rho_col.double().contiguous())
        return f.Analyse()

    def rmse(self):
        return float(((self.X.mean(0) - self.truth[0]) ** 2).mean().sqrt())

    def clone_state(self):
        return self.X.clone(), self.truth.clone(), self.gen.get_state()

    def set_state(self, st):
        X, t, g = st
        self.X = X.clone()
        self.truth = t.clone()
        self.gen.set_state(g)


def _gc(z):
    z = abs(z)
    if z >= 2.0:
        return 0.0
    if z <= 1.0:
        return (((-0.25 * z + 0.5) * z + 0.625) * z - 5.0 / 3.0) * z * z + 1.0
    return ((((z / 12.0 - 0.5) * z + 0.625) * z + 5.0 / 3.0) * z - 5.0) * z + 4.0 - 2.0 / (3.0 * z)


def cycle(twin: L96Twin, ncycles: int, inflator: Inflator | None = None, collect=False,
          log_every=0, progress=None, rho: float | None = None):
    """Forecast, observe, (features, target), analyse; returns the log and,
    with collect, the stacked features and targets of every cycle.  `rho`
    overrides the setup's fixed inflation.  A filter whose ensemble turns
    non-finite (catastrophic divergence of the cycled model) stops there:
    the library reports it (MDA_E_NUMERIC) and the log records the cycle in
    `diverged_at`, with an infinite RMSE for the remaining cycles."""
    torch = _torch()
    cfg = twin.config(rho)
    lg = CycleLog()
    feats, targets = [], []
    for c in range(ncycles):
        twin.forecast(twin.truth)
        twin.forecast(twin.X)
        y = twin.observe()
        rho_col = None
        try:
            if collect or inflator is not None:
                f = twin.features(y, cfg)
                if collect:
                    feats.append(f)
                    targets.append(twin.target_rho(twin.X).float())
                if inflator is not None:
                    rho_col = inflator.predict(f)
                    lg.rho_mean.append(float(rho_col.mean()))
            twin.analyse(y, cfg, rho_col)
        except L.MdaError as e:
            if e.code != L.MDA_E_NUMERIC:
                raise
            lg.diverged_at = c
        if lg.diverged_at is None and not math.isfinite(twin.rmse()):
            lg.diverged_at = c
        if lg.diverged_at is not None:
            if collect and feats:  # the diverging cycle's inputs are not training data
                feats.pop()
                targets.pop()
            lg.rmse += [math.inf] * (ncycles - c)
            lg.spread += [math.inf] * (ncycles - c)
            if progress:
                progress(f"filter diverged at cycle {c + 1}")
            break
        lg.rmse.append(twin.rmse())
        lg.spread.append(float(twin.X.var(0, unbiased=True).mean().sqrt()))
        if progress and log_every and (c + 1) % log_every == 0:
            progress(f"cycle {c + 1}/{ncycles}: rmse {lg.rmse[-1]:.4f} spread {lg.spread[-1]:.4f}")
    if collect:
        if not feats:
            return lg, None, None
        return lg, torch.cat(feats), torch.cat(targets)
    return lg


def _mean(v):
    return float(np.mean(v)) if len(v) else math.nan


def experiment(setup: L96Setup | None = None, train_cycles=1000, eval_cycles=1000,
               burn_in=100, train_steps=3000, val_cycles=200, train_rhos=None,
               fixed_rhos=None, ctx: Context | None = None, progress=None):
    """The twin-experiment protocol.

    Training data: the plain LETKF cycled with fixed inflation for
    `train_cycles` cycles from one initial state, once per rho in
    `train_rhos` (default: the setup's rho), the samples of cycles after
    `burn_in`.  The network is trained on all of them with the library's MLP
    kernels.  Then, for every fixed rho in `fixed_rhos` (default: the setup's
    rho), from the state of that rho's plain filter at cycle `train_cycles`:
      validation  cycles train_cycles + 1 .. + val_cycles, fixed and learned
                  side by side from the same state: the learned inflation is
                  deployed only if its analysis RMSE against the truth
                  (Metrics.hpp:265-273) is not worse there, else fixed rho;
      evaluation  the remaining eval_cycles - val_cycles cycles, each filter
                  continuing from its own state; the deployed filter's RMSE
                  is the reported result (the choice never sees these cycles).
    Returns one record per fixed rho."""
    torch = _torch()
    s = setup or L96Setup()
    ctx = ctx or Context(0)
    train_rhos = list(train_rhos or [s.rho_fixed])
    fixed_rhos = list(fixed_rhos or [s.rho_fixed])
    twin = L96Twin(ctx, s)
    st0 = twin.clone_state()
    feats, targets, runs, end_state = [], [], {}, {}
    for r in sorted(set(train_rhos) | set(fixed_rhos)):
        twin.set_state(st0)
        collect = r in train_rhos
        if progress:
            progress(f"plain filter rho={r}: {train_cycles} cycles" + (" (training data)" if collect else ""))
        out = cycle(twin, train_cycles, collect=collect, log_every=max(1, train_cycles // 4),
                    progress=progress, rho=r)
        lg = out[0] if collect else out
        runs[r] = lg
        end_state[r] = None if lg.diverged_at is not None else twin.clone_state()
        if collect and out[1] is not None:
            nb = burn_in * s.N
            if out[1].shape[0] > nb:
                feats.append(out[1][nb:])
                targets.append(out[2][nb:])
    if not feats:
        raise RuntimeError("no training data: every plain filter diverged during burn-in")
    feat, target = torch.cat(feats), torch.cat(targets)
    inf = Inflator(ctx, seed=s.seed)
    hist = inf.train(feat, target, steps=train_steps)
    if progress:
        progress(f"trained on {feat.shape[0]} samples: loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    results = []
    for r in fixed_rhos:
        rec = {"setup": s.__dict__ | {"radius_km": s.radius_km, "rho_fixed": r},
               "train_rhos": train_rhos, "train_cycles": train_cycles,
               "eval_cycles": eval_cycles, "val_cycles": val_cycles, "burn_in": burn_in,
               "train_samples": int(feat.shape[0]), "train_loss_first": hist[0],
               "train_loss_last": hist[-1],
               "rmse_train_fixed": _mean(runs[r].rmse[burn_in:]),
               "train_diverged_at": runs[r].diverged_at}
        if end_state[r] is None:
            rec.update({"rmse_fixed": math.inf, "rmse_learned": None, "fallback": None,
                        "rmse_deployed": None, "gain": None})
            results.append(rec)
            continue
        ev = eval_cycles - val_cycles
        twin.set_state(end_state[r])
        fv = cycle(twin, val_cycles, rho=r)
        st_fixed = twin.clone_state()
        twin.set_state(end_state[r])
        lv = cycle(twin, val_cycles, inflator=inf, rho=r)
        st_learn = twin.clone_state()
        twin.set_state(st_fixed)
        fe = cycle(twin, ev, rho=r)
        if lv.diverged_at is None:
            twin.set_state(st_learn)
            le = cycle(twin, ev, inflator=inf, rho=r)
        else:
            le = CycleLog(rmse=[math.inf] * ev, diverged_at=0)
        use_learned = _mean(lv.rmse) <= _mean(fv.rmse)
        r_fixed, r_learn = _mean(fe.rmse), _mean(le.rmse)
        deployed = r_learn if use_learned else r_fixed
        rec.update({
            "rmse_val_fixed": _mean(fv.rmse), "rmse_val_learned": _mean(lv.rmse),
            "rmse_fixed": r_fixed, "rmse_learned": r_learn,
            "spread_fixed": _mean(fe.spread), "spread_learned": _mean(le.spread),
            "rho_learned_mean": _mean(lv.rho_mean + le.rho_mean),
            "learned_diverged_at": lv.diverged_at if lv.diverged_at is not None else le.diverged_at,
            "fallback": not use_learned, "rmse_deployed": deployed,
            "gain": (r_fixed - deployed) / r_fixed,
            "gain_learned": (r_fixed - r_learn) / r_fixed})
        if progress:
            progress(f"rho={r}: eval rmse fixed {r_fixed:.4f} learned {r_learn:.4f} "
                     f"(validation {rec['rmse_val_fixed']:.4f} / {rec['rmse_val_learned']:.4f}, "
                     f"{'learned' if use_learned else 'fixed'} deployed)")
        results.append(rec)
    return results


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--N", type=int, default=4000)
    ap.add_argument("--k", type=int, default=40)
    ap.add_argument("--train-cycles", type=int, default=1000)
    ap.add_argument("--eval-cycles", type=int, default=1000)
    ap.add_argument("--val-cycles", type=int, default=200)
    ap.add_argument("--rho", type=float, nargs="+", default=[1.05])
    ap.add_argument("--train-rho", type=float, nargs="+", default=None)
    ap.add_argument("--support", type=float, default=16.0)
    ap.add_argument("--steps", type=int, default=3000)
    a = ap.parse_args()
    log = lambda m: print(f"[learned] {m}", file=sys.stderr, flush=True)
    res = experiment(L96Setup(N=a.N, k=a.k, rho_fixed=a.rho[0], support_gp=a.support),
                     train_cycles=a.train_cycles, eval_cycles=a.eval_cycles,
                     val_cycles=a.val_cycles, train_steps=a.steps,
                     train_rhos=a.train_rho or a.rho, fixed_rhos=a.rho, progress=log)
    print(json.dumps(res), flush=True)
