"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The product package (metada_amd) never imports
this module.

Array conventions follow oracle.h: ensembles are member-major float64 arrays
of shape (k, n); observation-space ensembles (k, p); small matrices (n, n)
row-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = C.c_double
_i64 = C.c_int64
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)


def build(force: bool = False) -> str:
    """Compile oracle.c with the committed Makefile (gcc, no GPU)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class EnkfDiag(C.Structure):
    _fields_ = [("innovation_norm", _d), ("analysis_increment_norm", _d),
                ("background_spread", _d), ("analysis_spread", _d),
                ("max_kalman_gain", _d), ("min_kalman_gain", _d),
                ("condition_number", _d), ("ensemble_size", C.c_int),
                ("observation_count", _i64)]


class LwenkfDiag(C.Structure):
    _fields_ = [("innovation_norm", _d), ("background_spread", _d), ("analysis_spread", _d),
                ("max_kalman_gain", _d), ("min_kalman_gain", _d), ("condition_number", _d),
                ("max_weight", _d), ("min_weight", _d), ("weight_variance", _d)]


LW_LOC = {"gaussian": 0, "exponential": 1, "cutoff": 2, "gaspari_cohn": 3}
LW_WEIGHT = {"uniform": 0, "adaptive": 1, "inverse_var": 2, "likelihood": 3}


class LetkfProblem(C.Structure):
    _fields_ = [("k", C.c_int), ("nx", _i64), ("ny", _i64), ("nlev", _i64),
                ("coord", C.c_int), ("col_a", _pd), ("col_b", _pd), ("col_c", _pd),
                ("p", _i64), ("obs_a", _pd), ("obs_b", _pd), ("obs_c", _pd),
                ("yo", _pd), ("var", _pd), ("inflation", _d), ("rho_col", _pd),
                ("radius", _d)]


class MetricsOut(C.Structure):
    _fields_ = [("rmse", _d), ("bias", _d), ("correlation", _d), ("crps", _d),
                ("avg_spread", _d)]


COORD_GRID, COORD_GEOGRAPHIC, COORD_CARTESIAN = 0, 1, 2
MLP_NPARAM = 8 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.orc_cholesky_lower.argtypes = [C.c_int, _pd, _pd]
        L.orc_lu_inverse.argtypes = [C.c_int, _pd, _pd]
        L.orc_jacobi_eigh.argtypes = [C.c_int, _pd, _pd, _pd, C.c_int, C.POINTER(C.c_int)]
        L.orc_singular_values.argtypes = [C.c_int, C.c_int, _pd, _pd]
        for f in ("orc_haversine_km", "orc_grid_distance"):
            getattr(L, f).restype = _d
            getattr(L, f).argtypes = [_d] * 4
        L.orc_cart_distance.restype = _d
        L.orc_cart_distance.argtypes = [_d] * 6
        L.orc_gaspari_cohn.restype = _d
        L.orc_gaspari_cohn.argtypes = [_d]
        L.orc_idw4.restype = _d
        L.orc_idw4.argtypes = [_pd, _i64, _i64, _i64, _d, _d, _d]
        L.orc_idw4_stencil.argtypes = [_i64, _i64, _i64, _d, _d, _d, _pi64, _pd]
        L.orc_csr_apply.argtypes = [_i64, _pi64, _pi64, _pd, _pd, _pd]
        L.orc_ensemble_mean.argtypes = [C.c_int, _i64, _pd, _pd]
        L.orc_ensemble_perturbations.argtypes = [C.c_int, _i64, _pd, _pd, _d, _pd]
        L.orc_enkf_analyse.argtypes = [C.c_int, _i64, _i64, _pd, _pd, _pd, _pd, _pd, _d,
                                       C.POINTER(EnkfDiag)]
        L.orc_etkf_analyse.argtypes = [C.c_int, _i64, _i64, _pd, _pd, _pd, _pd, _d]
        L.orc_lw_taper.restype = _d
        L.orc_lw_taper.argtypes = [_d, _d, C.c_int, C.c_int]
        L.orc_lwenkf_analyse.argtypes = [C.c_int, _i64, _i64, _pd, _pd, _pd, _pd, _pd, _d, _d,
                                         C.c_int, C.c_int, C.c_int, _pd, _pd, C.c_int,
                                         C.POINTER(LwenkfDiag)]
        L.orc_letkf_standard.argtypes = [C.POINTER(LetkfProblem), _pd, _pd, _pi64, _i64, _pd,
                                         _pi64, C.c_int]
        L.orc_letkf_refcompat_snapshot.argtypes = [C.POINTER(LetkfProblem), _pd, _pd, _pd]
        L.orc_letkf_refcompat_inplace.argtypes = [C.POINTER(LetkfProblem), _pd, _pi64, _pi64,
                                                  _pd]
        L.orc_metrics.argtypes = [C.c_int, _i64, _pd, _pd, _pd, _pd, C.POINTER(MetricsOut)]
        L.orc_l96_tendency.argtypes = [_i64, _pd, _d, _pd]
        L.orc_l96_euler_step.argtypes = [_i64, _pd, _d, _d]
        L.orc_l96_rk4_step.argtypes = [_i64, _pd, _d, _d]
        L.orc_mlp_forward.argtypes = [_i64, _pd, _pd, _pd]
        L.orc_mlp_loss_grad.argtypes = [_i64, _pd, _pd, _pd, _pd, _pd]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


def _pi(a):
    return a.ctypes.data_as(_pi64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")
    return rc


# ---- linear algebra ---------------------------------------------------------
def cholesky_lower(A):
    A = _f64(A); n = A.shape[0]; L = np.zeros_like(A)
    _check(lib().orc_cholesky_lower(n, _p(A), _p(L)), "cholesky")
    return L


def lu_inverse(A):
    A = _f64(A); n = A.shape[0]; Ai = np.zeros_like(A)
    _check(lib().orc_lu_inverse(n, _p(A), _p(Ai)), "lu_inverse")
    return Ai


def jacobi_eigh(A, max_sweeps=100):
    A = _f64(A); n = A.shape[0]
    w = np.zeros(n); V = np.zeros_like(A); sw = C.c_int(0)
    _check(lib().orc_jacobi_eigh(n, _p(A), _p(w), _p(V), max_sweeps, C.byref(sw)), "jacobi")
    return w, V, sw.value


def singular_values(A):
    A = _f64(A); m, n = A.shape; s = np.zeros(n)
    _check(lib().orc_singular_values(m, n, _p(A), _p(s)), "svd")
    return s


# ---- distances, localisation, H ---------------------------------------------
def haversine_km(lat1, lon1, lat2, lon2):
    return lib().orc_haversine_km(lat1, lon1, lat2, lon2)


def grid_distance(x1, y1, x2, y2):
    return lib().orc_grid_distance(x1, y1, x2, y2)


def cart_distance(x1, y1, z1, x2, y2, z2):
    return lib().orc_cart_distance(x1, y1, z1, x2, y2, z2)


def gaspari_cohn(z):
    return lib().orc_gaspari_cohn(float(z))


def idw4(state, nx, ny, nz, x, y, z):
    state = _f64(state)
    return lib().orc_idw4(_p(state), nx, ny, nz, x, y, z)


def idw4_stencil(nx, ny, nz, x, y, z):
    idx = np.zeros(4, np.int64); w = np.zeros(4)
    _check(lib().orc_idw4_stencil(nx, ny, nz, x, y, z, _pi(idx), _p(w)), "stencil")
    return idx, w


def csr_apply(row_ptr, col, w, state):
    row_ptr = np.ascontiguousarray(row_ptr, np.int64); col = np.ascontiguousarray(col, np.int64)
    w = _f64(w); state = _f64(state); p = len(row_ptr) - 1; out = np.zeros(p)
    lib().orc_csr_apply(p, _pi(row_ptr), _pi(col), _p(w), _p(state), _p(out))
    return out


# ---- ensemble statistics ------------------------------------------------------
def ensemble_mean(X):
    X = _f64(X); k, n = X.shape; m = np.zeros(n)
    lib().orc_ensemble_mean(k, n, _p(X), _p(m))
    return m


def ensemble_perturbations(X, mean, scale=1.0):
    X = _f64(X); mean = _f64(mean); k, n = X.shape; Xp = np.zeros_like(X)
    lib().orc_ensemble_perturbations(k, n, _p(X), _p(mean), scale, _p(Xp))
    return Xp


# ---- EnKF / ETKF --------------------------------------------------------------
def enkf_analyse(X, Yb, yo, var, Z, inflation):
    """Returns (X_a, diag dict). Z is (p, k)."""
    X = _f64(X).copy(); Yb = _f64(Yb); yo = _f64(yo); var = _f64(var); Z = _f64(Z)
    k, n = X.shape; p = Yb.shape[1]
    dg = EnkfDiag()
    _check(lib().orc_enkf_analyse(k, n, p, _p(X), _p(Yb), _p(yo), _p(var), _p(Z), inflation,
                                  C.byref(dg)), "enkf")
    return X, {f: getattr(dg, f) for f, _ in EnkfDiag._fields_}


def lw_taper(dist, radius, fn="gaussian", gc_standard=False):
    return lib().orc_lw_taper(float(dist), float(radius), LW_LOC[fn], int(gc_standard))


def lwenkf_analyse(X, Yb, yo, var, Z, inflation, radius, fn="gaussian", weighting="uniform",
                   gc_standard=False, want_K=False, want_cond=True):
    """Returns (X_a, weights, diag dict, K or None). Z is (p, k), Yb (k, p)."""
    X = _f64(X).copy(); Yb = _f64(Yb); yo = _f64(yo); var = _f64(var); Z = _f64(Z)
    k, n = X.shape; p = Yb.shape[1]
    w = np.zeros(k)
    K = np.zeros((n, p)) if want_K else None
    dg = LwenkfDiag()
    _check(lib().orc_lwenkf_analyse(k, n, p, _p(X), _p(Yb), _p(yo), _p(var), _p(Z), inflation,
                                    radius, LW_LOC[fn], int(gc_standard), LW_WEIGHT[weighting],
                                    _p(w), _p(K) if want_K else None, int(want_cond),
                                    C.byref(dg)), "lwenkf")
    return X, w, {f: getattr(dg, f) for f, _ in LwenkfDiag._fields_}, K


def etkf_analyse(X, Yb, yo, var, inflation):
    X = _f64(X).copy(); Yb = _f64(Yb); yo = _f64(yo); var = _f64(var)
    k, n = X.shape; p = Yb.shape[1]
    _check(lib().orc_etkf_analyse(k, n, p, _p(X), _p(Yb), _p(yo), _p(var), inflation), "etkf")
    return X


# ---- LETKF --------------------------------------------------------------------
class Problem:
    """Keeps the numpy arrays alive behind a LetkfProblem struct."""

    def __init__(self, k, nx, ny, nlev, coord, col_a, col_b, obs_a, obs_b, yo, var,
                 inflation, radius, rho_col=None, col_c=None, obs_c=None):
        self.arrays = {}

        def keep(name, a):
            if a is None:
                return None
            a = _f64(a)
            self.arrays[name] = a
            return _p(a)
        self.s = LetkfProblem()
        s = self.s
        s.k, s.nx, s.ny, s.nlev, s.coord = k, nx, ny, nlev, coord
        s.col_a, s.col_b, s.col_c = keep("col_a", col_a), keep("col_b", col_b), keep("col_c", col_c)
        s.p = len(yo)
        s.obs_a, s.obs_b, s.obs_c = keep("obs_a", obs_a), keep("obs_b", obs_b), keep("obs_c", obs_c)
        s.yo, s.var = keep("yo", yo), keep("var", var)
        s.inflation, s.radius = inflation, radius
        s.rho_col = keep("rho_col", rho_col)


def letkf_standard(prob: Problem, X, Yb, cols=None, want_T=False, threads=0):
    X = _f64(X).copy(); Yb = _f64(Yb)
    k = prob.s.k
    if cols is not None:
        cols = np.ascontiguousarray(cols, np.int64); nsel = len(cols)
    else:
        nsel = prob.s.nx * prob.s.ny
    T = np.zeros((nsel, k, k)) if want_T else None
    nloc = np.zeros(nsel, np.int64)
    _check(lib().orc_letkf_standard(C.byref(prob.s), _p(X), _p(Yb),
                                    _pi(cols) if cols is not None else None,
                                    nsel, _p(T) if want_T else None, _pi(nloc), threads),
           "letkf_standard")
    return X, T, nloc


def letkf_refcompat_snapshot(prob: Problem, X, Yb, want_t=False):
    X = _f64(X).copy(); Yb = _f64(Yb)
    t = np.zeros((prob.s.nx * prob.s.ny, prob.s.k)) if want_t else None
    _check(lib().orc_letkf_refcompat_snapshot(C.byref(prob.s), _p(X), _p(Yb),
                                              _p(t) if want_t else None), "letkf_compat")
    return (X, t) if want_t else X


def letkf_refcompat_inplace(prob: Problem, X, row_ptr, col, w):
    X = _f64(X).copy()
    row_ptr = np.ascontiguousarray(row_ptr, np.int64); col = np.ascontiguousarray(col, np.int64)
    w = _f64(w)
    _check(lib().orc_letkf_refcompat_inplace(C.byref(prob.s), _p(X), _pi(row_ptr), _pi(col),
                                             _p(w)), "letkf_inplace")
    return X


# ---- metrics, L96, MLP ----------------------------------------------------------
def metrics(X, truth):
    X = _f64(X); truth = _f64(truth); k, n = X.shape
    mean = np.zeros(n); spread = np.zeros(n); out = MetricsOut()
    _check(lib().orc_metrics(k, n, _p(X), _p(truth), _p(mean), _p(spread), C.byref(out)),
           "metrics")
    d = {f: getattr(out, f) for f, _ in MetricsOut._fields_}
    d["mean"], d["spread"] = mean, spread
    return d


def l96_tendency(x, F=8.0):
    x = _f64(x); d = np.zeros_like(x)
    lib().orc_l96_tendency(len(x), _p(x), F, _p(d))
    return d


def l96_step(x, F, dt, scheme="rk4"):
    x = _f64(x).copy()
    if scheme == "rk4":
        lib().orc_l96_rk4_step(len(x), _p(x), F, dt)
    else:
        lib().orc_l96_euler_step(len(x), _p(x), F, dt)
    return x


def mlp_forward(params, feat):
    params = _f64(params); feat = _f64(feat); m = feat.shape[0]; out = np.zeros(m)
    lib().orc_mlp_forward(m, _p(params), _p(feat), _p(out))
    return out


def mlp_loss_grad(params, feat, target):
    params = _f64(params); feat = _f64(feat); target = _f64(target)
    m = feat.shape[0]; loss = C.c_double(0); g = np.zeros(MLP_NPARAM)
    lib().orc_mlp_loss_grad(m, _p(params), _p(feat), _p(target), C.byref(loss), _p(g))
    return loss.value, g
