#!/usr/bin/env python3
"""LETKF analysis benchmark on one MI355X.

Workload (the metric's configuration): global regular lat-lon grid
1024 x 512 columns x 40 levels (n = 20,971,520), k = 64 members, Gaussian
random-field perturbations (200 km horizontal, 3-level vertical correlation),
p = 1e6 area-uniform point observations with trilinear H and sigma = 0.5,
Gaspari-Cohn localisation with 500 km support, rho = 1.05.

A step is one full LETKF analysis of the device-resident ensemble (H x and
innovations, local-observation bucketing, per-column Gram, eigensolve and
transform, in-place update).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm_GBs": 8000.0, "fp64_TFs": 78.6, "fp32_TFs": 157.3, "bf16_TFs": 2500.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--skip-plain", action="store_true",
                    help="skip the FP64 Jacobi (plain) path")
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="only copy this many GiB device-to-device once (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--batch", action="store_true",
                    help="batching form instead of the headline run: clustered observations "
                         "(50 Gaussian blobs of 300 km over a 10%% uniform floor, --p default 2e6), "
                         "analyses with k in --batch-k in one batch (hipGraph, work-queue Gram, "
                         "autotune) vs the same analyses one at a time")
    ap.add_argument("--batch-k", default="32,64,128")
    ap.add_argument("--learned", action="store_true",
                    help="learned-inflation form instead of the headline run: cycled Lorenz-96 "
                         "(N=4000, k=40, 1000 training + 1000 evaluation cycles), the network "
                         "trained on the plain filter's cycles at every --learned-rho, then "
                         "learned vs fixed inflation at each of them")
    ap.add_argument("--learned-rho", default="1.02,1.05,1.10")
    ap.add_argument("--lwenkf", action="store_true",
                    help="LWEnKF form instead of the headline run: a 2048 x 2048 state "
                         "(n = 4,194,304), k = 64, p = 512 IDW4 observations, adaptive weights, "
                         "gaussian index-distance taper of radius 0.05, rho = 1.05")
    ap.add_argument("--transform-precision", default="fp64", choices=["fp64", "bf16x3"],
                    help="precision of X_a = xbar + X'T on the timed path (the split-precision "
                         "form is also measured beside it, with its errors, in the full run)")
    ap.add_argument("--value-path-only", action="store_true",
                    help="run only the --precision path (for profiler runs): no oracle, "
                         "no FP64 reference, no error report")
    ap.add_argument("--nlon", type=int, default=1024)
    ap.add_argument("--nlat", type=int, default=512)
    ap.add_argument("--nlev", type=int, default=40)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--p", type=int, default=1_000_000)
    ap.add_argument("--radius", type=float, default=500.0)
    ap.add_argument("--rho", type=float, default=1.05)
    ap.add_argument("--sample-cols", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads", type=int, default=0)
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {a.gpus} is not supported in this build "
              f"(the LETKF path shards by columns across GPUs only in design).", file=sys.stderr)
        return 2
    log("start: importing torch")
    import numpy as np
    import torch
    log(f"torch {torch.__version__}, device {torch.cuda.get_device_name(0)}")
    if a.copy_calibration_gib > 0:
        nel = int(a.copy_calibration_gib * (1 << 30)) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nel * 8}), flush=True)
        return 0
    import metada_amd as mda
    from metada_amd import synthetic
    from metada_amd import _lib as L

    dev = "cuda:0"
    if a.batch:
        return run_batch(a, log)
    if a.learned:
        return run_learned(a, log)
    if a.lwenkf:
        return run_lwenkf(a, log)
    t0 = time.time()
    W = synthetic.latlon_workload(a.nlon, a.nlat, a.nlev, a.k, a.p, device=dev, progress=log)
    torch.cuda.synchronize()
    log(f"workload generated in {time.time() - t0:.1f}s")
    ncol, k, nlev, p = a.nlon * a.nlat, a.k, a.nlev, a.p
    n = ncol * nlev

    # ---- sampled-column oracle (parity at full size and the CPU baseline) ----
    rng = np.random.default_rng(17)
    cols = np.sort(rng.choice(ncol, size=min(a.sample_cols, ncol), replace=False))
    col_lat = W["geom"].col_a.cpu().numpy()
    col_lon = W["geom"].col_b.cpu().numpy()
    idx = (np.arange(nlev)[:, None] * ncol + cols[None, :]).ravel()       # [lev][sample]
    if not a.value_path_only:
        Xs = W["X"][:, torch.from_numpy(idx).to(dev)].cpu().numpy()
        hcols = W["H"].col.view(p, 8)
        Yb = np.empty((k, p))
        for j in range(k):  # H x_j member by member (bounded temporaries)
            Yb[j] = (W["X"][j][hcols] * W["H"].weight.view(p, 8)).sum(-1).cpu().numpy()
        log("oracle inputs staged on the host")
    yo = W["obs"].values.cpu().numpy()
    var = W["obs"].variances.cpu().numpy()
    rho32 = float(np.float32(a.rho))
    cpu = None
    T_oracle = None
    try:
        if a.no_cpu_baseline or a.value_path_only:
            raise RuntimeError("skipped (--no-cpu-baseline)")
        from oracle import oracle as orc
        threads = a.threads or len(os.sched_getaffinity(0))
        prob = orc.Problem(k, len(cols), 1, nlev, orc.COORD_GEOGRAPHIC, col_lat[cols],
                           col_lon[cols], W["lat"], W["lon"], yo, var, rho32,
                           float(np.float32(a.radius)))
        t1 = time.time()
        orc.letkf_standard(orc.Problem(k, 1, 1, nlev, orc.COORD_GEOGRAPHIC, col_lat[cols[:1]],
                                       col_lon[cols[:1]], W["lat"], W["lon"], yo, var, rho32,
                                       float(np.float32(a.radius))),
                           Xs[:, ::len(cols)].copy(), Yb, threads=threads)
        t_fixed = time.time() - t1
        t1 = time.time()
        Xo, T_oracle, nloc_o = orc.letkf_standard(prob, Xs, Yb, want_T=True, threads=threads)
        t_sample = time.time() - t1
        per_col = max(t_sample - t_fixed, 1e-9) / len(cols)
        cpu_s = t_fixed + per_col * ncol
        # one thread on the first 32 sampled columns (the reference is single-threaded)
        n1 = min(32, len(cols))
        sel = (np.arange(nlev)[:, None] * len(cols) + np.arange(n1)[None, :]).ravel()
        prob1 = orc.Problem(k, n1, 1, nlev, orc.COORD_GEOGRAPHIC, col_lat[cols[:n1]],
                            col_lon[cols[:n1]], W["lat"], W["lon"], yo, var, rho32,
                            float(np.float32(a.radius)))
        t1 = time.time()
        orc.letkf_standard(orc.Problem(k, 1, 1, nlev, orc.COORD_GEOGRAPHIC, col_lat[cols[:1]],
                                       col_lon[cols[:1]], W["lat"], W["lon"], yo, var, rho32,
                                       float(np.float32(a.radius))),
                           Xs[:, ::len(cols)].copy(), Yb, threads=1)
        t_fixed1 = time.time() - t1
        t1 = time.time()
        orc.letkf_standard(prob1, Xs[:, sel].copy(), Yb, threads=1)
        t_one = time.time() - t1
        cpu1_s = t_fixed1 + max(t_one - t_fixed1, 1e-9) / n1 * ncol
        cpu_model = ""
        try:
            with open("/proc/cpuinfo") as fh:
                cpu_model = next((ln.split(":", 1)[1].strip() for ln in fh
                                  if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": 1.0 / cpu_s, "unit": "analyses/s", "cores": threads, "kind": "port",
               "sample": f"{len(cols)} of {ncol} columns (all {nlev} levels, full local-obs "
                         f"search over p={p}), oracle C FP64 standard LETKF with OpenMP; "
                         f"{t_sample:.2f}s for the sample + {t_fixed:.2f}s fixed, "
                         f"extrapolated to {cpu_s:.0f}s per analysis",
               "value_1_thread": 1.0 / cpu1_s,
               "sample_1_thread": f"{n1} columns on 1 thread in {t_one:.2f}s ({t_fixed1:.2f}s "
                                  f"fixed), extrapolated to {cpu1_s:.0f}s per analysis",
               "cpu_model": cpu_model, "nproc": os.cpu_count()}
        log(f"oracle sample: {t_sample:.2f}s ({threads} threads), mean local obs "
            f"{nloc_o.mean():.0f}")
    except Exception as e:  # the oracle is a checker; the bench still runs without it
        log(f"oracle unavailable: {e!r}")

    # ---- GPU analyses ----
    # Every analysis starts from the same background X_b (restored outside the
    # timed region), so the paths are timed on, and compared over, one input.
    ctx = mda.Context(0)
    Xb = W["X"].clone()
    cfg = mda.Config({"inflation": a.rho, "localization_radius": a.radius})
    ens = mda.Ensemble(W["X"], a.nlon, a.nlat, nlev)

    tp_value = {"fp64": L.PREC_FP64, "bf16x3": L.PREC_BF16X3}[a.transform_precision]

    def make(prec_code, eig, tp=L.PREC_FP64, keep=False):
        return mda.LETKF(ens, W["obs"], W["H"], cfg, geometry=W["geom"], ctx=ctx,
                         precision=prec_code, eig_method=eig, transform_precision=tp,
                         keep_transforms=keep)

    def run(name, f, warmup, steps, check_parity=False):
        times, diags = [], []
        for it in range(warmup + steps):
            W["X"].copy_(Xb)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            d = f.Analyse()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t1) * 1000.0
            if it == 0 and check_parity and T_oracle is not None:
                Tg = f.transforms(cols)
                err = float(np.max(np.abs(Tg - T_oracle)) / np.max(np.abs(T_oracle)))
                nonlocal_parity[name] = err
                log(f"{name}: max rel err of T vs the oracle on {len(cols)} columns = {err:.3e}")
            if it == 0 and d["n_fallback_columns"]:
                fl = f.column_flags()
                log(f"{name}: fallback reasons: " + ", ".join(
                    f"{nm}={int(((fl & bit) != 0).sum())}" for bit, nm in
                    ((1, "ql"), (2, "residual"), (4, "orthogonality"), (8, "below_sigma"))))
            tag = "warmup" if it < warmup else "step"
            log(f"{name} {tag} {it if it < warmup else it - warmup}: {dt:.1f} ms (lib "
                f"{d['ms_total']:.1f}: obs {d['ms_obs']:.1f}, bucket {d['ms_bucket']:.1f}, gram "
                f"{d['ms_gram']:.1f}, solve {d['ms_solve']:.1f} [trd {d['ms_eig_tridiag']:.1f}, "
                f"vals {d['ms_eig_values']:.1f}, vecs {d['ms_eig_vectors']:.1f}, jacobi "
                f"{d['ms_solve_jacobi']:.1f}], transform {d['ms_transform']:.1f}; fallback "
                f"{d['n_fallback_columns']})")
            if it >= warmup:
                times.append(dt)
                diags.append(d)
        return np.array(times), diags

    nonlocal_parity = {}
    stage_keys = ("ms_obs", "ms_bucket", "ms_gram", "ms_solve", "ms_eig_tridiag",
                  "ms_eig_values", "ms_eig_vectors", "ms_solve_jacobi", "ms_transform")

    def summary(times, diags):
        st = {key: float(np.mean([d[key] for d in diags])) for key in stage_keys}
        return {"ms_median": float(np.median(times)), "ms_min": float(np.min(times)),
                "ms_mean": float(np.mean(times)), "analyses_per_s": 1000.0 / float(np.median(times)),
                "stages_ms": st, "fallback_columns": int(diags[-1]["n_fallback_columns"])}

    if a.value_path_only:
        code = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32, "bf16x3": L.PREC_BF16X3}[a.precision]
        tv, dv = run(a.precision, make(code, L.EIG_AUTO, tp_value), a.warmup, a.steps)
        sv = summary(tv, dv)
        print(json.dumps({"metric": "profiler run", "precision": a.precision,
                          "ms_median": sv["ms_median"], "stages_ms": sv["stages_ms"]}), flush=True)
        ctx.close()
        return 0
    # FP64 with every column's T kept (eigenvectors and transform unfused):
    # T of the sampled columns against the oracle
    fk = make(L.PREC_FP64, L.EIG_AUTO, keep=True)
    tk, dk = run("fp64_kept", fk, 0, 1, check_parity=True)
    Xk = W["X"].clone()
    # FP64, tridiagonal eigensolver with the fused transform (fastest FP64
    # path): reference output
    f64 = make(L.PREC_FP64, L.EIG_AUTO)
    t64, d64 = run("fp64", f64, 1, max(1, min(a.steps, 5)))
    Xd = W["X"].clone()
    res = {"fp64": summary(t64, d64)}
    res["fp64_kept"] = summary(tk, dk)
    res["fp64"]["max_abs_diff_vs_kept"] = float((Xd - Xk).abs().max())
    log(f"fp64: fused vs kept-T path max |diff| = {res['fp64']['max_abs_diff_vs_kept']:.3e}")
    del Xk
    # FP64 plain: FP64 contractions and the batched Jacobi eigensolver
    if not a.skip_plain:
        fpl = make(L.PREC_FP64, L.EIG_JACOBI)
        tpl, dpl = run("fp64_jacobi", fpl, 1, max(1, min(a.steps, 3)), check_parity=True)
        res["fp64_jacobi"] = summary(tpl, dpl)
        res["fp64_jacobi"]["max_abs_diff_vs_fp64"] = float((W["X"] - Xd).abs().max())

    def errors(Xm):
        diff = Xm - Xd
        inc = Xd - Xb
        rms = lambda t: float(torch.sqrt(torch.mean(t * t)))
        mean_m, mean_d, mean_b = Xm.mean(0), Xd.mean(0), Xb.mean(0)
        return {"ensemble": {"rel_rms": rms(diff) / rms(Xd),
                             "rel_max": float(diff.abs().max() / Xd.abs().max()),
                             "rms_over_increment": rms(diff) / rms(inc)},
                "mean": {"rel_rms": rms(mean_m - mean_d) / rms(mean_d),
                         "rms_over_increment": rms(mean_m - mean_d) / rms(mean_d - mean_b)},
                "perturbations": {"rel_rms": rms((Xm - mean_m) - (Xd - mean_d)) / rms(Xd - mean_d)}}

    mixed = {}
    for pname in ("fp32", "bf16x3"):
        code = {"fp32": L.PREC_FP32, "bf16x3": L.PREC_BF16X3}[pname]
        is_value = pname == a.precision
        fm = make(code, L.EIG_AUTO, tp_value if is_value else L.PREC_FP64)
        tm, dm = run(pname, fm, a.warmup if is_value else 1,
                     a.steps if is_value else max(1, min(a.steps, 3)))
        mixed[pname] = summary(tm, dm)
        mixed[pname]["error_vs_fp64"] = errors(W["X"])
        log(f"{pname}: errors vs FP64 {json.dumps(mixed[pname]['error_vs_fp64'])}")
    res.update(mixed)
    # the precision form with X'W split-precision too (BF16x3 Gram and transform)
    if a.transform_precision == "fp64" and k % 16 == 0 and k <= 64:
        fx = make(L.PREC_BF16X3, L.EIG_AUTO, L.PREC_BF16X3)
        tx, dx = run("bf16x3_xw", fx, 1, max(1, min(a.steps, 3)))
        res["bf16x3_xw"] = summary(tx, dx)
        res["bf16x3_xw"]["error_vs_fp64"] = errors(W["X"])
        log(f"bf16x3_xw: errors vs FP64 {json.dumps(res['bf16x3_xw']['error_vs_fp64'])}")
    if a.precision == "fp64":
        f64b = make(L.PREC_FP64, L.EIG_AUTO)
        tv, dv = run("fp64_value", f64b, a.warmup, a.steps)
        vres = summary(tv, dv)
        d_last = dv[-1]
    else:
        vres = res[a.precision]
        d_last = None
    torch.cuda.synchronize()
    ms = vres["ms_mean"]
    value = 1000.0 / ms
    stages = vres["stages_ms"]
    d = d_last or f64.diag

    # ---- roofline of the dominant kernel (algorithmic work / HIP-event time) ----
    # Work per launch, priced by what each kernel computes:
    #   Gram        2 sum_g p_g k^2 + 2 sum_g p_g k  (C on the matrix cores, b)
    #   E1          (4/3) k^3 per column             (Householder reduction)
    #   E2          30 k^2 per column                (QL sweeps, ~2 per eigenvalue)
    #   E3          4 k^3 + 4 k^2 per column         (back-transform 2 k^3, W 2 k^3,
    #                                                 wbar; twisted factorisation O(k^2))
    #   transform   2 n k 8 bytes                    (X read once, X_a written once)
    #   obs stats   p (8 taps k 8 + 8 taps 16) + p k (8 + 4 if FP32 rows) + p 16 bytes
    tot_loc = d["total_local_obs"]
    nonempty = ncol - d["n_empty_columns"]
    gram_flop = 2.0 * tot_loc * k * k + 2.0 * tot_loc * k
    xform_bytes = 2.0 * n * k * 8
    obs_bytes = p * (8 * k * 8 + 8 * 16) + p * k * (8 + (4 if a.precision != "fp64" else 0)) + p * 16
    peak_c = {"fp64": PEAK["fp64_TFs"], "fp32": PEAK["fp32_TFs"], "bf16x3": PEAK["bf16_TFs"]}
    gram_kernel = "k_letkf_gram_f64" if a.precision == "fp64" else "k_letkf_gram_mixed"
    e1 = nonempty * (4.0 / 3.0) * k ** 3
    e2 = nonempty * 30.0 * k * k
    e3 = nonempty * (4.0 * k ** 3 + 4.0 * k * k)
    # with an FP64 transform at k in {16, 32, 48, 64} the eigenvector kernel
    # applies X_a = X T' itself (2 n k FLOP, X in and out); the transform
    # stage then covers only the columns left to the Jacobi fallback
    fused = tp_value == L.PREC_FP64 and k in (16, 32, 48, 64)
    if fused:
        e3 += 2.0 * n * k
    # (kernel, bound, achieved, peak, unit, ms, what limits it)
    cand = {
        "gram": (gram_kernel, "mfma", gram_flop / (stages["ms_gram"] * 1e-3) / 1e12,
                 peak_c[a.precision], "TFLOP/s", stages["ms_gram"],
                 "VALU operand scaling and hi/lo split" if a.precision == "bf16x3" else
                 "matrix-core issue"),
        "eig_tridiag": ("k_trd", "fp64", e1 / max(stages["ms_eig_tridiag"], 1e-9) / 1e9,
                        PEAK["fp64_TFs"], "TFLOP/s", stages["ms_eig_tridiag"],
                        "FP64 VALU issue and latency (matrix-vector products, rank-2 updates)"),
        "eig_values": ("k_trd_eigvals", "fp64", e2 / max(stages["ms_eig_values"], 1e-9) / 1e9,
                       PEAK["fp64_TFs"], "TFLOP/s", stages["ms_eig_values"],
                       "FP64 VALU latency (serial QL recurrence per lane)"),
        "eig_vectors": ("k_trd_vectors_x" if fused else "k_trd_vectors", "fp64",
                        e3 / max(stages["ms_eig_vectors"], 1e-9) / 1e9,
                        PEAK["fp64_TFs"], "TFLOP/s", stages["ms_eig_vectors"],
                        "FP64 VALU issue and latency at one wave per SIMD" +
                        ("; X update in 32-byte column segments with the CU's four "
                         "wavefronts waiting on it" if fused else "")),
    }
    if not fused:
        cand["transform"] = (
            ("k_letkf_transform_bf16x3" if tp_value == L.PREC_BF16X3 else
             "k_letkf_transform_mfma") if k % 16 == 0 and k <= 64 else
            "k_letkf_transform_reg" if k in (8, 20, 40) else "k_letkf_transform", "hbm",
            xform_bytes / (stages["ms_transform"] * 1e-3) / 1e9,
            PEAK["hbm_GBs"], "GB/s", stages["ms_transform"],
            "HBM access pattern (member-major X in 32-byte column segments)")
    dom = max(cand, key=lambda kk: cand[kk][5])
    kern, b, ach, pk, unit, _, limiter = cand[dom]
    # HBM bytes per launch of that kernel from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs (scaled by a copy of known size), for the same workload
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_letkf_{a.precision}_{a.nlon}x{a.nlat}x{nlev}_k{k}.json")
    if os.path.exists(tf) and p == 1_000_000:
        with open(tf) as fh:
            tj = json.load(fh)
        for name, v in tj["kernels"].items():
            if name.split("::")[-1].split("<")[0] == kern:
                traffic = v["traffic"]
    # the contract's two roofs: HBM bandwidth, or the arithmetic peak (the FP64
    # eigensolver kernels run on the vector units, whose FP64 peak is the same
    # 78.6 TFLOP/s; the stage entries below name the unit)
    roof = {"kernel": kern, "bound": "hbm" if b == "hbm" else "mfma", "achieved": ach,
            "peak": pk, "unit": unit, "frac": ach / pk, "traffic": traffic,
            "ms_per_launch": cand[dom][5], "limiter": limiter}
    # every stage against its own roof (the streaming stages in HBM GB/s)
    stage_roofs = {key: {"kernel": v[0], "bound": v[1], "achieved": v[2], "peak": v[3],
                         "unit": v[4], "frac": v[2] / v[3], "ms": v[5]}
                   for key, v in cand.items()}
    if fused:
        # the X bytes the eigenvector kernel moves, against the HBM roof
        stage_roofs["eig_vectors"]["hbm_GBs_algorithmic"] = (
            xform_bytes / (stages["ms_eig_vectors"] * 1e-3) / 1e9)
        stage_roofs["transform_fallback_columns"] = {"ms": stages["ms_transform"]}
    stage_roofs["obs_stats"] = {
        "kernel": "k_obs_stats (stage: tap-order sort + statistics)", "bound": "hbm",
        "achieved": obs_bytes / (stages["ms_obs"] * 1e-3) / 1e9, "peak": PEAK["hbm_GBs"],
        "unit": "GB/s", "frac": obs_bytes / (stages["ms_obs"] * 1e-3) / 1e9 / PEAK["hbm_GBs"],
        "ms": stages["ms_obs"]}
    ratio_jacobi = (res["fp64_jacobi"]["ms_median"] / vres["ms_median"]
                    if "fp64_jacobi" in res else None)
    out = {
        "metric": "LETKF analyses/sec at n=2.1e7,k=64: mixed-prec vs FP64 plain; RMS err vs FP64",
        "value": value, "unit": "analyses/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16x3": "bf16x3+f64"}[a.precision],
        "data": "synthetic (seeded Gaussian random fields, uniform obs)",
        "config": {"workload": f"LETKF global lat-lon {a.nlon}x{a.nlat}x{nlev}, k={k}, "
                               f"p={p}, trilinear H, GC {a.radius:g} km, rho={a.rho}",
                   "n": n, "k": k, "p": p, "precision": a.precision,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": stages,
        "stage_roofline": stage_roofs,
        # like for like: the plain FP64 path with the same eigensolver and
        # transform, only the contraction precision differs
        "speedup_vs_fp64_plain": res["fp64"]["ms_median"] / vres["ms_median"],
        # against the FP64 path whose eigensolver is the batched Jacobi kernel
        "speedup_vs_fp64_jacobi_eigensolver": ratio_jacobi,
        "paths": res,
        "local_obs": {"total": tot_loc, "mean": tot_loc / ncol, "max": d["max_local_obs"],
                      "empty_columns": d["n_empty_columns"]},
        "parity": {"oracle": "standard LETKF (C, FP64)", "columns": len(cols),
                   "max_rel_err_T": nonlocal_parity} if nonlocal_parity else None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def run_batch(a, log):
    """Batching form: the same grid with clustered observations, one analysis
    per ensemble size in --batch-k, run (1) one at a time and (2) as one batch
    (concurrent branches of one hipGraph, work-queue Gram, autotuned launch)."""
    import numpy as np
    import torch
    import metada_amd as mda
    from metada_amd import synthetic
    from metada_amd import _lib as L
    ks = [int(x) for x in a.batch_k.split(",")]
    p = a.p if a.p != 1_000_000 else 2_000_000
    codes = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32, "bf16x3": L.PREC_BF16X3}
    precs = [a.precision for k in ks]
    # the one-at-a-time runs use a context of their own, closed (workspace
    # freed) before the batch allocates its per-analysis buffers
    ctx = mda.Context(0)
    work = []
    for i, k in enumerate(ks):
        W = synthetic.latlon_workload(a.nlon, a.nlat, a.nlev, k, p, seed=4000 + i, clustered=True,
                                      device="cuda:0", progress=log)
        work.append((k, W, W["X"].clone()))
        log(f"workload k={k} generated")
    cfg = mda.Config({"inflation": a.rho, "localization_radius": a.radius})

    def filters():
        return [mda.LETKF(mda.Ensemble(W["X"], a.nlon, a.nlat, a.nlev), W["obs"], W["H"], cfg,
                          geometry=W["geom"], ctx=ctx, precision=codes[pr])
                for (_, W, _), pr in zip(work, precs)]

    def restore():
        for _, W, xb in work:
            W["X"].copy_(xb)
        torch.cuda.synchronize()

    # one at a time
    seq_ms, ref, hist = [], [], None
    for it in range(a.warmup + a.steps):
        restore()
        t1 = time.perf_counter()
        ds = [f.Analyse() for f in filters()]
        torch.cuda.synchronize()
        if it >= a.warmup:
            seq_ms.append((time.perf_counter() - t1) * 1e3)
        if it == 0:
            ref = [W["X"].clone() for _, W, _ in work]
            fs = filters()
            counts = []
            for f in fs:
                restore()
                f.Analyse()
                counts.append(f.local_counts())
            hist = [{"k": k, "max": int(c.max()), "mean": float(c.mean()),
                     "empty_columns": int((c == 0).sum()),
                     "percentiles": {q: float(np.percentile(c, q)) for q in (10, 50, 90, 99)}}
                    for (k, _, _), c in zip(work, counts)]
            log(f"local obs: {hist}")
        log(f"sequential {it}: {sum(d['ms_total'] for d in ds):.1f} ms (lib; per analysis "
            f"{[round(d['ms_total'], 1) for d in ds]})")
        for k, d in zip(ks, ds):
            log(f"  k={k}: " + ", ".join(f"{key[3:]} {d[key]:.1f}" for key in
                ("ms_obs", "ms_bucket", "ms_gram", "ms_eig_tridiag", "ms_eig_values",
                 "ms_eig_vectors", "ms_solve_jacobi", "ms_transform")))
    seq_stages = [{key: d[key] for key in d if key.startswith("ms_")} for d in ds]
    ctx.close()
    torch.cuda.empty_cache()
    ctx = mda.Context(0)
    # batched
    flags = L.BATCH_GRAPH | L.BATCH_QUEUE | L.BATCH_AUTOTUNE
    fb = filters()
    bat_ms, infos = [], []
    for it in range(a.warmup + a.steps):
        restore()
        t1 = time.perf_counter()
        diags, info = mda.letkf_batch(fb, flags)
        torch.cuda.synchronize()
        if it >= a.warmup:
            bat_ms.append((time.perf_counter() - t1) * 1e3)
        infos.append(info)
        if it == 0:
            same = all(torch.equal(W["X"], r) for (_, W, _), r in zip(work, ref))
            log(f"batch reproduces the sequential analyses bit for bit: {same}")
        log(f"batch {it}: {info['ms_total']:.1f} ms (graph replayed {info['graph_replayed']}, "
            f"gram blocks {info['gram_blocks']}, queue {info['gram_queue']})")
    out = {"metric": "batched LETKF analyses (clustered obs, mixed k) per second",
           "value": 1000.0 / float(np.median(bat_ms)), "unit": "batches/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": float(np.median(bat_ms)), "higher_is_better": True,
           "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16x3": "bf16x3+f64"}[a.precision],
           "data": "synthetic (GRF backgrounds, 50 Gaussian blobs of 300 km + 10% uniform)",
           "config": {"workload": f"batched LETKF {a.nlon}x{a.nlat}x{a.nlev}, k={ks}, p={p} "
                                  f"clustered, GC {a.radius:g} km", "k": ks, "p": p,
                      "precision": dict(zip(map(str, ks), precs))},
           "sequential_ms_median": float(np.median(seq_ms)),
           "batched_ms_median": float(np.median(bat_ms)),
           "speedup_batched_vs_sequential": float(np.median(seq_ms) / np.median(bat_ms)),
           "autotuned": {"gram_blocks": infos[-1]["gram_blocks"], "gram_queue": infos[-1]["gram_queue"]},
           "local_obs": hist,
           "sequential_stage_ms": dict(zip(map(str, ks), seq_stages))}
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def run_learned(a, log):
    """Learned-inflation form: the twin-experiment protocol of
    metada_amd.learned at its full size, one record per fixed rho."""
    from metada_amd import learned
    rhos = [float(x) for x in a.learned_rho.split(",")]
    t0 = time.time()
    res = learned.experiment(learned.L96Setup(N=4000, k=40, rho_fixed=rhos[0]),
                             train_cycles=1000, eval_cycles=1000, val_cycles=200, burn_in=100,
                             train_steps=3000, train_rhos=rhos, fixed_rhos=rhos,
                             progress=log)
    wall = time.time() - t0
    rows = [{"rho_fixed": r["setup"]["rho_fixed"], "rmse_fixed": r["rmse_fixed"],
             "rmse_learned": r["rmse_learned"], "rmse_deployed": r["rmse_deployed"],
             "deployed": "fixed" if r["fallback"] else "learned", "gain": r["gain"],
             "rmse_val_fixed": r.get("rmse_val_fixed"),
             "rmse_val_learned": r.get("rmse_val_learned"),
             "spread_fixed": r.get("spread_fixed"), "spread_learned": r.get("spread_learned"),
             "rho_learned_mean": r.get("rho_learned_mean")} for r in res]
    out = {"metric": "learned vs fixed inflation: analysis RMSE vs truth (cycles 1201-2000)",
           "value": min(r["rmse_deployed"] for r in res), "unit": "RMSE",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f64 filter, f32 MLP",
           "data": "synthetic (Lorenz-96 twin experiment, seeded)",
           "config": {"workload": "cycled LETKF on Lorenz-96 N=4000, k=40, dt=0.05 RK4, every "
                                  "2nd variable observed, sigma=1, GC support 16 grid points",
                      "train_cycles": 1000, "validation_cycles": 200, "eval_cycles": 800,
                      "train_rhos": rhos},
           "train_samples": res[0]["train_samples"], "train_loss": [res[0]["train_loss_first"],
                                                                    res[0]["train_loss_last"]],
           "wall_s": wall, "results": rows}
    print(json.dumps(out), flush=True)
    return 0


def run_lwenkf(a, log):
    """LWEnKF form (LWEnKF.hpp:207-334): one step is one full analysis of the
    device-resident ensemble.  The roofline is that of the fused gain and
    update kernel, k_lw_update: 4 n p k FLOP per launch (the K tile
    X'_rho G and its product with D), FP64 on the vector ALUs."""
    import numpy as np
    import torch
    import metada_amd as mda
    from metada_amd import _lib as L
    nx, ny, k, p = 2048, 2048, 64, 512
    n = nx * ny
    dev = "cuda:0"
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    xs = torch.linspace(0.0, 6.283185307179586, nx, device=dev, dtype=torch.float64)
    truth = (torch.sin(3.0 * xs)[None, :] * torch.cos(2.0 * xs)[:, None]).reshape(-1)
    X = truth[None, :] + 0.5 * torch.randn(k, n, generator=gen, device=dev, dtype=torch.float64)
    rng = np.random.default_rng(3)
    ox, oy = rng.uniform(0, nx - 1, p), rng.uniform(0, ny - 1, p)
    var = np.full(p, 0.25)
    ctx = mda.Context(0)
    ob0 = mda.Observation(np.zeros(p), var, ox, oy, np.zeros(p))
    H = mda.ObsOperator.identity_idw4(ob0, nx, ny, 1, ctx=ctx)
    t_host = truth.cpu().numpy()
    yo = mda.analysis.ObsOperator(H.row_ptr, H.col, H.weight).apply_host(t_host[None, :])[0] \
        + 0.5 * rng.standard_normal(p)
    ob = mda.Observation(yo, var, ox, oy, np.zeros(p))
    Z = torch.randn(p, k, generator=gen, device=dev, dtype=torch.float64)
    cfg = mda.Config({"analysis": {"inflation": a.rho, "localization_radius": 0.05,
                                   "inflation_method": "multiplicative",
                                   "localization_function": "gaussian",
                                   "weighting_scheme": "adaptive"}})
    X0 = X.clone()
    ens = mda.Ensemble(X, nx, ny, 1)
    f = mda.LWEnKF(ens, ob, H, cfg, ctx=ctx)
    upd, sol, wall = [], [], []
    for i in range(a.warmup + a.steps):
        X.copy_(X0)  # every step analyses the same background (outside the timed region)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.Analyse(Z=Z)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) * 1e3
        if i >= a.warmup:
            st = f.stage_ms()
            wall.append(dt)
            upd.append(st["ms_update"])
            sol.append(st["ms_gain_solve"])
            log(f"lwenkf step {i - a.warmup}: {dt:.1f} ms (update {st['ms_update']:.2f}, "
                f"S^-T Y' {st['ms_gain_solve']:.2f})")
    ms = float(np.median(wall))
    flops = 4.0 * n * p * k
    achieved = flops / (float(np.mean(upd)) * 1e-3) / 1e12
    res = f.getAnalysisResults()
    cpu = None
    if not a.no_cpu_baseline:
        # the oracle on the same p, k and weights with fewer state rows: its
        # fixed part (S, S^-1, cond(S)) at 256 rows and its per-row part
        # from 16384 rows, extrapolated to n rows; one core
        from oracle import oracle as orc
        X.copy_(X0)
        Xh = X[:, :16384].cpu().numpy()
        Yb = np.stack([mda.analysis.ObsOperator(H.row_ptr, H.col, H.weight)
                       .apply_host(X[j].cpu().numpy()[None, :])[0] for j in range(k)])
        Zh = Z.cpu().numpy()
        t1 = time.time()
        orc.lwenkf_analyse(Xh[:, :256], Yb, yo, var, Zh, f.inflation_factor,
                           f.localization_radius, "gaussian", "adaptive", want_cond=False)
        t_fix = time.time() - t1
        t1 = time.time()
        orc.lwenkf_analyse(Xh, Yb, yo, var, Zh, f.inflation_factor, f.localization_radius,
                           "gaussian", "adaptive", want_cond=False)
        t_big = time.time() - t1
        per_row = max(t_big - t_fix, 0.0) / (16384 - 256)
        t_full = t_fix + per_row * n
        cpu = {"value": 1.0 / t_full, "unit": "analyses/s", "cores": 1, "kind": "port",
               "sample": f"oracle (oracle/oracle.c) at p={p}, k={k}, without the cond(S) "
                         f"diagnostic (its one-sided Jacobi SVD is not a performance path): "
                         f"{t_fix:.2f} s at 256 state rows, {t_big:.2f} s at 16384 rows, "
                         f"extrapolated linearly in rows to n={n}"}
    out = {"metric": "LWEnKF analyses/sec", "value": 1e3 / ms, "unit": "analyses/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded random ensemble around a smooth field)",
           "config": {"workload": "LWEnKF 2048x2048 state (n=4194304), k=64, p=512 IDW4 obs, "
                                  "adaptive weights, gaussian index taper r=0.05, rho=1.05",
                      "n": n, "k": k, "p": p},
           "roofline": {"kernel": "k_lw_update", "bound": "fp64-valu", "achieved": achieved,
                        "peak": PEAK["fp64_TFs"], "unit": "TFLOP/s",
                        "frac": achieved / PEAK["fp64_TFs"], "traffic": None,
                        "ms_per_launch": float(np.mean(upd))},
           "stages_ms": {"update": float(np.mean(upd)), "gain_solve": float(np.mean(sol))},
           "cpu_baseline": cpu,
           "diagnostics": {key: res[key] for key in ("condition_number", "max_weight",
                                                     "min_weight", "max_kalman_gain",
                                                     "analysis_spread")}}
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
